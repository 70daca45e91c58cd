// csg_device.h - device records and HIP kernels of the gravity core (gfx950).
// Included by csg_core.hip only (one translation unit holds every kernel).
//
// Pipeline per substep (all on one stream, no host synchronisation):
//   bbox  : min/max over live bodies -> root cube            (src/octree.c:119-159)
//   keys  : 16-level octant path per body from the FP64 cell recurrence (48 bits)
//                                                            (src/octree.c:30-50, 87-108)
//   sort  : LSD radix sort of (key, body) pairs, 6 passes of 8 bits
//   runs  : bodies sharing all 16 levels get their levels 16..63 and are
//           re-ranked inside the run                         (src/octree.c:71-79 depth limit)
//   lcp   : common-prefix depth of sorted neighbours -> nodes per body
//   scan  : node ids in depth-first pre-order
//   dd    : double-double prefix sums of m, m*x, m*y, m*z over sorted bodies
//   emit  : one 64-byte walk record per node (COM, mass, (2h)^2, skip, leaf body)
//   walk  : one wavefront per 64 sorted bodies, stackless pre-order traversal
//           with each lane applying its own opening test   (src/octree.c:171-218)
//   kdk   : kick / drift                                    (src/integrator.c:22-31, 45-51)
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "body.h"
#include "octree.h"

#define CSG_TPB 256
#define CSG_ITEMS 8
#define CSG_TILE (CSG_TPB * CSG_ITEMS)
// Levels in the sorted key.  Bodies sharing all of them form collision runs
// resolved with their deeper levels (k_runs; 8 runs on the 1M-body merger, whose
// leaves sit at depths 10-18).  3 * CSG_KEY_LEVELS bits, sorted in CSG_SORT_PASSES
// 8-bit passes.  13 levels take one pass fewer, but leave ~4,000 runs to resolve:
// the 1M merger's build took 0.366 against 0.348 ms.  A live key is never all ones
// in those bits (that cell lies in the 1% margin of the root cube), so the
// all-ones key of dead bodies sorts last.
#define CSG_KEY_LEVELS 16
// Node records past the last node: padding the wave-mask walks may step onto
// (they test for the end once per round of eight FP32 or four FP64 steps);
// never accepted, no mass
#define CSG_NODE_PAD 8
#define CSG_KEY_BITS (3 * CSG_KEY_LEVELS)
#define CSG_SORT_PASSES ((CSG_KEY_BITS + 7) / 8)
// collision words: ext[0] levels L..L+20, ext[1] L+21..L+41, ext[2] L+42..63
#define CSG_EXT2_LEVELS (64 - CSG_KEY_LEVELS - 42)
static_assert(CSG_EXT2_LEVELS >= 1 && CSG_EXT2_LEVELS <= 21, "three collision words cover levels L..63");
#define CSG_DEPTH_LIMIT 64
#define CSG_BBOX_BLOCKS 1024

// Device-resident build state, written by the kernels themselves.
struct BuildParams {
    double root[4];  // centre x, y, z and half size H of the root cube
    int n_live;      // bodies with !(mass <= 0)
    int n_nodes;     // nodes of the tree (reference count: internal + leaves)
    int n_runs;      // groups of sorted bodies with equal keys (all CSG_KEY_LEVELS levels)
    int max_depth;   // deepest leaf
    int overflow;    // 1 when n_nodes exceeds the node capacity
    int node_cap;
    int sort_stall;  // 1 when a radix pass gave up waiting for an earlier tile
    int fail_count;  // builds that overflowed or stalled before the current one (never
                     // reset by a build: a failure in the middle of a substep run stays
                     // visible to the check at the run's end)
    int order_bad;   // 1 when a Peano-Hilbert rank fell outside [0, n_live)
    unsigned long long interactions;  // body-node interactions of the last walk
};

// One tree node in depth-first pre-order.  skip = index of the first node after
// this node's subtree, so "descend" is i+1 and "accept / leaf" is skip.
struct __align__(64) WNode {
    double cx, cy, cz;  // centre of mass
    double m;           // total mass
    double s2;          // (2h)^2 with h the cell half size
    int skip;
    int body;           // leaf: original body index (first one for a depth-limit leaf); else -1
    int depth;
    int first;          // member range [first, last] in sorted order
    int last;
    int pad;
};
static_assert(sizeof(WNode) == 64, "WNode is one 64-byte record");

// The same node for the FP32 walk (40 bytes).  The centre of mass stays FP64:
// a lane forms its body-node offset as fp32(C - x) with the subtraction in FP64
// (full rate on CDNA4), so the offset carries a single FP32 rounding however far
// the pair sits from the origin.  Mass is FP32; s2 is (2h)^2 / theta^2 for the
// opening angle the records were emitted for (divided in FP64, +inf at theta 0),
// so the walk's opening test is s2 < d^2.  Softening stays out of the test: with
// eps^2 on both sides, FP32 resolves the criterion of nodes much smaller than
// eps only to eps^2's rounding, and the opening decisions drift from the FP64
// walk's.
struct __align__(8) WNode32 {
    double cx, cy, cz;  // centre of mass
    float m;            // total mass
    float s2;           // (2h)^2 / theta^2; -1 for a leaf
    int skip;
    int body;
};
static_assert(sizeof(WNode32) == 40, "WNode32 is one 40-byte record");

// Dispatch order of the walks: wave slot s walks wave order[s] (s < slots,
// order[s] >= 0); cost (may be null) receives each wave's node-step count.
struct WaveOrder {
    const int *order;
    int n;        // waves
    int *cost;
    int slots;    // entries of order (slots past it, or holding -1, walk nothing)
};

// One record of a walk tape: a node visited by a wavefront and the lanes
// (bodies) that interacted with it there.  A tape is the frozen list of
// interactions of one force evaluation, in each wave's visiting order.
struct TapeRec {
    unsigned long long imask;
    int node;
    int pad;
};
static_assert(sizeof(TapeRec) == 16, "TapeRec is 16 bytes");

// Where a walk writes its tape: wave w owns rec[w*cap, w*cap + count[w]).
struct TapeOut {
    TapeRec *rec;
    int *count;
    int cap;
    int *overflow;
};

__device__ __forceinline__ void tape_push(const TapeOut &t, int wave, int lane, int &nrec,
                                          unsigned long long imask, int node) {
    if (imask == 0ull) return;
    if (lane == 0 && nrec < t.cap) {
        TapeRec r;
        r.imask = imask;
        r.node = node;
        r.pad = 0;
        t.rec[(size_t)wave * t.cap + nrec] = r;
    }
    ++nrec;
}

__device__ __forceinline__ void tape_close(const TapeOut &t, int wave, int lane, int nrec) {
    if (lane != 0) return;
    t.count[wave] = nrec < t.cap ? nrec : t.cap;
    if (nrec > t.cap) atomicOr(t.overflow, 1);
}

__device__ __forceinline__ void store_nodes(WNode *__restrict__ nodes, WNode32 *__restrict__ n32,
                                            int id, const WNode &w, double inv_th2) {
    if (nodes) nodes[id] = w;   // FP64 records: skipped when only the FP32 walk reads the tree
    WNode32 f;
    f.cx = w.cx;
    f.cy = w.cy;
    f.cz = w.cz;
    f.m = (float)w.m;
    f.s2 = w.body >= 0 ? -1.0f : (float)(w.s2 * inv_th2);  // a leaf passes every test
    f.skip = w.skip;
    f.body = w.body;
    n32[id] = f;
}

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ bool live_mass(double m) { return !(m <= 0.0); }

// 1/sqrt(r2) in FP64: hardware estimate refined by two Newton steps on the
// residual e = 1 - r2 y^2 (each step squares the relative error).
__device__ __forceinline__ double rsq64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    double e = fma(-r2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-r2 * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// The FP64 walks' 1/sqrt(r2): the hardware estimate refined by one Newton step
// on the same residual.
__device__ __forceinline__ double rsq64_walk(double r2) {
    const double y = __builtin_amdgcn_rsq(r2);
    const double e = fma(-r2 * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// Octant code: bit0 x >= c, bit1 y < c, bit2 z < c.
__device__ __forceinline__ int cell_octant(double px, double py, double pz, double cx, double cy,
                                           double cz) {
    return (px >= cx ? 1 : 0) | (py < cy ? 2 : 0) | (pz < cz ? 4 : 0);
}

// Child cell of octant q: centre moves by half the parent half size (the
// product is exact), half size halves.
__device__ __forceinline__ void cell_child(int q, double &cx, double &cy, double &cz, double &h) {
    const double o = 0.5 * h;
    cx = (q & 1) ? cx + o : cx - o;
    cy = (q & 2) ? cy - o : cy + o;
    cz = (q & 4) ? cz - o : cz + o;
    h = h * 0.5;
}

// `levels` octant codes of point p starting from cell (c, h), 3 bits each,
// first level in the most significant position; (c, h) advance.
__device__ __forceinline__ uint64_t path_word(double px, double py, double pz, double &cx,
                                              double &cy, double &cz, double &h, int levels) {
    uint64_t k = 0;
    for (int l = 0; l < levels; ++l) {
        const int q = cell_octant(px, py, pz, cx, cy, cz);
        k = (k << 3) | (uint64_t)q;
        cell_child(q, cx, cy, cz, h);
    }
    return k;
}

// Levels CSG_KEY_LEVELS..63 of a body path in three words (21, 21 and
// CSG_EXT2_LEVELS levels).
__device__ __forceinline__ void path_ext(double px, double py, double pz, const double *root,
                                         uint64_t *e) {
    double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    (void)path_word(px, py, pz, cx, cy, cz, h, CSG_KEY_LEVELS);
    e[0] = path_word(px, py, pz, cx, cy, cz, h, 21);
    e[1] = path_word(px, py, pz, cx, cy, cz, h, 21);
    e[2] = path_word(px, py, pz, cx, cy, cz, h, CSG_EXT2_LEVELS);
}

// Number of leading levels two sorted bodies share (0..64).  Bodies with equal
// keys belong to one collision run and have their ext words.  (A word of w
// levels sits in the low 3w bits, so its first level is bit 3w-1.)
__device__ __forceinline__ int lcp_sorted(const uint64_t *__restrict__ key,
                                          const uint64_t *__restrict__ ext, int a, int b) {
    constexpr int L = CSG_KEY_LEVELS;
    const uint64_t ka = key[a], kb = key[b];
    if (ka != kb) return (__clzll((long long)(ka ^ kb)) - (64 - 3 * L)) / 3;
    const uint64_t *ea = ext + 3 * (size_t)a, *eb = ext + 3 * (size_t)b;
    if (ea[0] != eb[0]) return L + (__clzll((long long)(ea[0] ^ eb[0])) - 1) / 3;
    if (ea[1] != eb[1]) return L + 21 + (__clzll((long long)(ea[1] ^ eb[1])) - 1) / 3;
    if (ea[2] != eb[2])
        return L + 42 + (__clzll((long long)(ea[2] ^ eb[2])) - (64 - 3 * CSG_EXT2_LEVELS)) / 3;
    return 64;
}

// Octant digit of sorted body b at level l.
__device__ __forceinline__ int digit_at(const uint64_t *key, const uint64_t *ext, int b, int l) {
    constexpr int L = CSG_KEY_LEVELS;
    if (l < L) return (int)((key[b] >> (3 * (L - 1 - l))) & 7);
    const uint64_t *e = ext + 3 * (size_t)b;
    if (l < L + 21) return (int)((e[0] >> (3 * (L + 20 - l))) & 7);
    if (l < L + 42) return (int)((e[1] >> (3 * (L + 41 - l))) & 7);
    return (int)((e[2] >> (3 * (63 - l))) & 7);
}

// Last sorted body j >= b sharing at least d levels with b (galloping search).
__device__ int range_end(const uint64_t *key, const uint64_t *ext, int b, int d, int n_live) {
    int good = b, step = 1;
    while (true) {
        const int j = good + step;
        if (j >= n_live || lcp_sorted(key, ext, b, j) < d) break;
        good = j;
        step <<= 1;
    }
    int bad = min(good + step, n_live);
    while (bad - good > 1) {
        const int mid = good + ((bad - good) >> 1);
        if (lcp_sorted(key, ext, b, mid) >= d) good = mid; else bad = mid;
    }
    return good;
}

// First sorted body j <= b sharing at least d levels with b.
__device__ int range_begin(const uint64_t *key, const uint64_t *ext, int b, int d) {
    int good = b, step = 1;
    while (true) {
        const int j = good - step;
        if (j < 0 || lcp_sorted(key, ext, j, b) < d) break;
        good = j;
        step <<= 1;
    }
    int bad = max(good - step, -1);
    while (good - bad > 1) {
        const int mid = bad + ((good - bad + 1) >> 1);
        if (lcp_sorted(key, ext, mid, b) >= d) good = mid; else bad = mid;
    }
    return good;
}

// ---------------------------------------------------------- double-double
struct DD { double h, l; };

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    double s, e, t, f;
    two_sum(a.h, b.h, s, e);
    two_sum(a.l, b.l, t, f);
    e += t;
    double s2 = s + e;
    double e2 = e - (s2 - s);
    e2 += f;
    DD r;
    r.h = s2 + e2;
    r.l = e2 - (r.h - s2);
    return r;
}
__device__ __forceinline__ DD dd_neg(DD a) { return DD{-a.h, -a.l}; }
__device__ __forceinline__ DD dd_prod(double a, double b) {
    DD r;
    r.h = a * b;
    r.l = fma(a, b, -r.h);
    return r;
}

// --------------------------------------------------------- block helpers
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one int per thread over a 256-thread block.
__device__ __forceinline__ int block_excl_scan(int v, int *lds4, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int inc = wave_incl_scan(v);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int s = lds4[k];
        if (k < w) off += s;
        total += s;
    }
    __syncthreads();
    return off + inc - v;
}

// Control block of the body sort: the tile tickets of the passes and the count
// of every 8-bit digit value in every pass (pass p sorts bits 8p..8p+7).
#define RS_TPB 1024                  // threads per block of the sort passes
#define RS_ITEMS 4                   // keys per thread in the sort passes
#define RS_WAVES (RS_TPB / 64)
#define RS_TILE (RS_TPB * RS_ITEMS)  // keys per tile
struct RsCtl {
    int ticket[CSG_SORT_PASSES];
    int hist[CSG_SORT_PASSES][256];
};

// =================================================================== bbox
__global__ __launch_bounds__(CSG_TPB) void k_bbox_partial(const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          const double *__restrict__ z,
                                                          const double *__restrict__ m, int n,
                                                          double *__restrict__ part) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double cnt = 0.0;
    for (int i = blockIdx.x * CSG_TPB + threadIdx.x; i < n; i += gridDim.x * CSG_TPB) {
        if (!live_mass(m[i])) continue;
        cnt += 1.0;
        const double p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
    __shared__ double s[7][CSG_TPB];
    for (int a = 0; a < 3; ++a) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
    s[6][threadIdx.x] = cnt;
    __syncthreads();
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            for (int a = 0; a < 3; ++a) {
                if (s[a][o] < s[a][threadIdx.x]) s[a][threadIdx.x] = s[a][o];
                if (s[3 + a][o] > s[3 + a][threadIdx.x]) s[3 + a][threadIdx.x] = s[3 + a][o];
            }
            s[6][threadIdx.x] += s[6][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x][0];
}

// Root cube exactly as src/octree.c:141-159; also resets the per-build counters.
__global__ __launch_bounds__(CSG_TPB) void k_bbox_final(const double *__restrict__ part, int nblk,
                                                        BuildParams *bp, int node_cap,
                                                        RsCtl *__restrict__ ctl) {
    // the sort's tickets and digit counts start from zero in every build
    for (int k = threadIdx.x; k < (int)(sizeof(RsCtl) / sizeof(int)); k += CSG_TPB)
        ((int *)ctl)[k] = 0;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double cnt = 0.0;
    for (int b = threadIdx.x; b < nblk; b += CSG_TPB) {
        for (int a = 0; a < 3; ++a) {
            const double l = part[b * 8 + a], h = part[b * 8 + 3 + a];
            if (l < lo[a]) lo[a] = l;
            if (h > hi[a]) hi[a] = h;
        }
        cnt += part[b * 8 + 6];
    }
    __shared__ double s[7][CSG_TPB];
    for (int a = 0; a < 3; ++a) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
    s[6][threadIdx.x] = cnt;
    __syncthreads();
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            for (int a = 0; a < 3; ++a) {
                if (s[a][o] < s[a][threadIdx.x]) s[a][threadIdx.x] = s[a][o];
                if (s[3 + a][o] > s[3 + a][threadIdx.x]) s[3 + a][threadIdx.x] = s[3 + a][o];
            }
            s[6][threadIdx.x] += s[6][o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int live = (int)s[6][0];
        // the previous build's failure is counted before its flags are cleared
        if (bp->overflow || bp->sort_stall || bp->order_bad) bp->fail_count += 1;
        bp->n_live = live;
        bp->n_runs = 0;
        bp->max_depth = 0;
        bp->overflow = 0;
        bp->sort_stall = 0;
        bp->order_bad = 0;
        bp->node_cap = node_cap;
        bp->interactions = 0ull;
        bp->n_nodes = 1;
        if (live == 0) {
            bp->root[0] = 0.0; bp->root[1] = 0.0; bp->root[2] = 0.0; bp->root[3] = 1.0;
        } else {
            double ext = s[3][0] - s[0][0];
            if (s[4][0] - s[1][0] > ext) ext = s[4][0] - s[1][0];
            if (s[5][0] - s[2][0] > ext) ext = s[5][0] - s[2][0];
            double half = ext * 0.5 * 1.01;
            if (half < 1e-10) half = 1.0;
            bp->root[0] = (s[0][0] + s[3][0]) * 0.5;
            bp->root[1] = (s[1][0] + s[4][0]) * 0.5;
            bp->root[2] = (s[2][0] + s[5][0]) * 0.5;
            bp->root[3] = half;
        }
    }
}

// =================================================================== keys
// Dead bodies get the all-ones key and sort behind every live body.
// Also writes (x, y, z, m) interleaved, one 32-byte line per body, so the sorted
// gather after the sort reads one sector per body instead of four, and counts
// the digit values of every sort pass (all passes' offsets from this one pass
// over the bodies): tiles of RS_TILE bodies, per-tile counts in LDS added to the
// sort's global counts.  Clears the npart tile counters of the node-count scan.
__global__ __launch_bounds__(RS_TPB) void k_keys(const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z,
                                                 const double *__restrict__ m, int n,
                                                 const BuildParams *__restrict__ bp,
                                                 uint64_t *__restrict__ key,
                                                 double4 *__restrict__ xyzm,
                                                 RsCtl *__restrict__ ctl,
                                                 int *__restrict__ part, int npart) {
    __shared__ int h[CSG_SORT_PASSES][256];
    for (int g = blockIdx.x * RS_TPB + threadIdx.x; g < npart; g += gridDim.x * RS_TPB) part[g] = 0;
    for (int k = threadIdx.x; k < CSG_SORT_PASSES * 256; k += RS_TPB) (&h[0][0])[k] = 0;
    __syncthreads();
    const double r0 = bp->root[0], r1 = bp->root[1], r2 = bp->root[2], r3 = bp->root[3];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = blockIdx.x * RS_TILE + r * RS_TPB + threadIdx.x;
        if (i >= n) break;
        uint64_t k = ~0ull;
        const double px = x[i], py = y[i], pz = z[i], pm = m[i];
        if (live_mass(pm)) {
            double cx = r0, cy = r1, cz = r2, hh = r3;
            k = path_word(px, py, pz, cx, cy, cz, hh, CSG_KEY_LEVELS);
        }
        key[i] = k;
        xyzm[i] = make_double4(px, py, pz, pm);
#pragma unroll
        for (int p = 0; p < CSG_SORT_PASSES; ++p) atomicAdd(&h[p][(int)((k >> (8 * p)) & 255)], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CSG_SORT_PASSES * 256; k += RS_TPB) {
        const int v = (&h[0][0])[k];
        if (v) atomicAdd(&(&ctl->hist[0][0])[k], v);
    }
}


// ============================================================ radix sort
// Tile of CSG_TILE consecutive elements per block; wave w owns the sub-tile
// [w*512, w*512+512) and walks it in 8 rounds of 64, so (wave, round, lane) is
// the element order and the per-wave counters make the scatter stable.  Keys
// are 64-bit (neighbour queries) or 32-bit (the adjoint's node ids).
template <typename K>
__global__ __launch_bounds__(CSG_TPB) void k_rs_hist(const K *__restrict__ key, int n,
                                                     int shift, int *__restrict__ hist, int nb) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * CSG_TILE;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int i = base + r * CSG_TPB + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of the digit's per-tile counts.
__global__ __launch_bounds__(CSG_TPB) void k_rs_rowscan(int *__restrict__ hist, int nb,
                                                        int *__restrict__ tot) {
    __shared__ int lds4[4];
    int *row = hist + (size_t)blockIdx.x * nb;
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += CSG_TPB) {
        const int i = c0 + threadIdx.x;
        const int v = i < nb ? row[i] : 0;
        int total;
        const int ex = block_excl_scan(v, lds4, total);
        if (i < nb) row[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <typename K>
__global__ __launch_bounds__(CSG_TPB) void k_rs_scatter(
    const K *__restrict__ kin, const int *__restrict__ vin, K *__restrict__ kout,
    int *__restrict__ vout, int n, int shift, const int *__restrict__ hist,
    const int *__restrict__ tot, int nb) {
    __shared__ int wcnt[4][256];
    __shared__ int binbase[256];
    __shared__ int lds4[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < 4 * 256; k += CSG_TPB) (&wcnt[0][0])[k] = 0;
    {
        int total;
        binbase[threadIdx.x] = block_excl_scan(tot[threadIdx.x], lds4, total);
    }
    __syncthreads();
    const int base = blockIdx.x * CSG_TILE + w * (64 * CSG_ITEMS);
    const uint64_t lt = (1ull << lane) - 1ull;
    K kk[CSG_ITEMS];
    int vv[CSG_ITEMS], dg[CSG_ITEMS], rk[CSG_ITEMS];
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int i = base + r * 64 + lane;
        const bool valid = i < n;
        kk[r] = valid ? kin[i] : (K)0;
        vv[r] = valid ? vin[i] : 0;
        const int d = (int)((kk[r] >> shift) & 255);
        dg[r] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1;
            const uint64_t bal = __ballot(on);
            peers &= on ? bal : ~bal;
        }
        const int below = __popcll(peers & lt);
        const int old = valid ? wcnt[w][d] : 0;
        if (valid && below == 0) wcnt[w][d] = old + __popcll(peers);
        rk[r] = old + below;
    }
    __syncthreads();
    // per-digit offsets of each wave inside the tile, and where each digit's
    // run starts in the tile's digit-sorted order
    __shared__ int woff[4][256];
    __shared__ int tstart[256];
    {
        const int d = threadIdx.x;
        int s = 0;
        for (int k = 0; k < 4; ++k) { woff[k][d] = s; s += wcnt[k][d]; }
        int total;
        tstart[d] = block_excl_scan(s, lds4, total);
    }
    __syncthreads();
    // stage the tile in digit-sorted order, then write each digit run with
    // consecutive lanes (coalesced)
    __shared__ K skey[CSG_TILE];
    __shared__ int sval[CSG_TILE];
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int i = base + r * 64 + lane;
        if (i < n) {
            const int d = dg[r];
            const int lp = tstart[d] + woff[w][d] + rk[r];
            skey[lp] = kk[r];
            sval[lp] = vv[r];
        }
    }
    __syncthreads();
    const int cnt = min(CSG_TILE, n - blockIdx.x * CSG_TILE);
    for (int j = threadIdx.x; j < cnt; j += CSG_TPB) {
        const K k = skey[j];
        const int d = (int)((k >> shift) & 255);
        const int pos = binbase[d] + hist[d * nb + blockIdx.x] + (j - tstart[d]);
        kout[pos] = k;
        vout[pos] = sval[j];
    }
}

// ============================================== single-sweep radix sort
// The body sort: CSG_SORT_PASSES LSD passes of 8-bit digits over the key
// bits, one kernel per pass.  The counts of every digit value come from k_keys, so a pass
// needs no histogram of its own; what a tile needs besides is how many keys of
// each digit value the tiles before it hold.  Tiles are numbered by ticket in
// the order their blocks start.  Each tile publishes its own count per digit
// value at once, and the running total (its own plus everything before) as
// soon as it knows it; a tile learns its prefix by adding published counts
// backwards from the previous tile until it meets a running total (decoupled
// look-back).  A tile only waits on tiles with smaller tickets, whose blocks
// are resident already, so every wait ends; the waits are also bounded
// (RS_SPIN_LIMIT rounds that find nothing new) after which the build flags a stalled sort (and an
// overflow, so the kernels after it do nothing) instead of waiting further.
//
// status[tile][digit] = epoch << 33 | running-total flag << 32 | count.  Every
// pass uses a new epoch, so words left by earlier passes read as not yet
// published and the array is never cleared between passes.
#define RS_SPIN_LIMIT (1 << 20)
#define RS_WINDOW 8

__device__ __forceinline__ uint64_t rs_word(unsigned epoch, bool total, int count) {
    return ((uint64_t)epoch << 33) | ((uint64_t)(total ? 1 : 0) << 32) | (uint64_t)(unsigned)count;
}

// Load the tile's keys and values: wave w holds elements w*64*RS_ITEMS + r*64 +
// lane (round r), the order the ranking keeps.
template <bool FIRST>
__device__ __forceinline__ void rs_load(const uint64_t *__restrict__ kin,
                                        const int *__restrict__ vin, int n, int tile, int w,
                                        int lane, uint64_t (&kk)[RS_ITEMS], int (&vv)[RS_ITEMS]) {
    const int base = tile * RS_TILE + w * (64 * RS_ITEMS) + lane;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int i = base + r * 64;
        kk[r] = i < n ? kin[i] : ~0ull;
        vv[r] = FIRST ? i : (i < n ? vin[i] : 0);
    }
}

__device__ __forceinline__ uint64_t rs_status_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pass `pass` (bits 8*pass .. 8*pass+7).  FIRST: the values are the positions
// of the keys (the body indices), so vin is not read.  Stable: within a tile
// the element order is (wave, round, lane), as the loads.
//
// Order of work, so the waits overlap: the tile is loaded speculatively as
// tile blockIdx.x before the ticket is taken (blocks almost always start in
// order; a different ticket reloads); the tile's digit counts (LDS adds) are
// published at once and the first look-back window is requested; the stable
// ranking (ballots, 16 waves per block so their latencies overlap) runs while
// those loads are in flight.  Threads 0..255 own one digit value each.
template <bool FIRST>
__global__ __launch_bounds__(RS_TPB) void k_rs_sweep(
    const uint64_t *__restrict__ kin, const int *__restrict__ vin, uint64_t *__restrict__ kout,
    int *__restrict__ vout, int n, int pass, RsCtl *__restrict__ ctl,
    uint64_t *__restrict__ status, unsigned epoch, BuildParams *__restrict__ bp) {
    __shared__ int s_tile;
    __shared__ int tcnt[256];
    __shared__ unsigned short wcnt[RS_WAVES][256];   // per-wave counts, then offsets
    __shared__ int tstart[256];
    __shared__ int binpos[256];
    __shared__ int lback[256];
    __shared__ uint64_t skey[RS_TILE];
    __shared__ int sval[RS_TILE];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool owner = tid < 256;   // owns digit value d = tid
    const int d = tid & 255;
    uint64_t kk[RS_ITEMS];
    int vv[RS_ITEMS];
    rs_load<FIRST>(kin, vin, n, (int)blockIdx.x, w, lane, kk, vv);
    if (tid == 0) s_tile = atomicAdd(&ctl->ticket[pass], 1);
    if (owner) tcnt[d] = 0;
    for (int k = tid; k < RS_WAVES * 256; k += RS_TPB) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const int tile = s_tile;
    if (tile != (int)blockIdx.x) rs_load<FIRST>(kin, vin, n, tile, w, lane, kk, vv);
    const int shift = 8 * pass;
    const int base = tile * RS_TILE + w * (64 * RS_ITEMS);
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r)
        if (base + r * 64 + lane < n) atomicAdd(&tcnt[(int)((kk[r] >> shift) & 255)], 1);
    __syncthreads();
    int cnt_d = 0, t = tile - 1;
    uint64_t v[RS_WINDOW];
    uint64_t *const mine = status + (size_t)tile * 256 + d;
    if (owner) {
        cnt_d = tcnt[d];
        __hip_atomic_store(mine, rs_word(epoch, tile == 0, cnt_d), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // first look-back window, consumed after the ranking
#pragma unroll
        for (int k = 0; k < RS_WINDOW; ++k)
            v[k] = t - k >= 0 ? rs_status_load(status + (size_t)(t - k) * 256 + d) : 0ull;
    }
    // rank inside the wave's sub-tile: lanes with the same digit value are
    // found with eight ballots (lane j matches bit b when ballot_b ^ flip_b has
    // bit j set, flip_b = all ones where this lane's bit b is clear); the lowest
    // such lane bumps the wave's counter
    const uint64_t lt = (1ull << lane) - 1ull;
    int rk[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const bool valid = base + r * 64 + lane < n;
        const int dg = (int)((kk[r] >> shift) & 255);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t bal = __ballot((dg >> bit) & 1);
            const uint64_t flip = (uint64_t)(int64_t)(((dg >> bit) & 1) - 1);
            peers &= bal ^ flip;
        }
        const int below = __popcll(peers & lt);
        const int old = valid ? (int)wcnt[w][dg] : 0;
        if (valid && below == 0) wcnt[w][dg] = (unsigned short)(old + __popcll(peers));
        rk[r] = old + below;
    }
    // look-back for digit value d, RS_WINDOW tiles per round trip: counts are
    // added from tile-1 downwards up to the first running total; an unpublished
    // word ends the round and is requested again
    if (owner) {
        int before = 0;
        if (tile > 0) {
            int polls = 0;
            for (;;) {
                bool done = false;
                int used = 0;
#pragma unroll
                for (int k = 0; k < RS_WINDOW; ++k) {
                    if (done || used < k || t - k < 0 || (unsigned)(v[k] >> 33) != epoch) continue;
                    before += (int)(unsigned)v[k];
                    used = k + 1;
                    done = ((v[k] >> 32) & 1ull) != 0;
                }
                if (done) break;
                t -= used;
                if (t < 0) break;   // not reached: tile 0 publishes a running total
                if (used == 0) {
                    if (++polls >= RS_SPIN_LIMIT) {
                        atomicOr(&bp->sort_stall, 1);
                        atomicOr(&bp->overflow, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
#pragma unroll
                for (int k = 0; k < RS_WINDOW; ++k)
                    v[k] = t - k >= 0 ? rs_status_load(status + (size_t)(t - k) * 256 + d) : 0ull;
            }
            __hip_atomic_store(mine, rs_word(epoch, true, before + cnt_d), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        lback[d] = before;
    }
    __syncthreads();
    if (owner) {   // per-wave counts -> each wave's offset inside the digit's run
        int o = 0;
#pragma unroll
        for (int k = 0; k < RS_WAVES; ++k) {
            const int c = wcnt[k][d];
            wcnt[k][d] = (unsigned short)o;
            o += c;
        }
    }
    if (w == 0) {
        // digit runs inside the tile and in the output: exclusive scans over the
        // 256 digit values, four per lane
        int c[4], g[4], sc = 0, sg = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = tcnt[4 * lane + j];
            g[j] = ctl->hist[pass][4 * lane + j];
            sc += c[j];
            sg += g[j];
        }
        int oc = wave_incl_scan(sc) - sc, og = wave_incl_scan(sg) - sg;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            tstart[4 * lane + j] = oc;
            binpos[4 * lane + j] = og + lback[4 * lane + j];
            oc += c[j];
            og += g[j];
        }
    }
    __syncthreads();
    // stage the tile in digit order, then each digit run leaves with
    // consecutive lanes
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (base + r * 64 + lane < n) {
            const int dg = (int)((kk[r] >> shift) & 255);
            const int lp = tstart[dg] + (int)wcnt[w][dg] + rk[r];
            skey[lp] = kk[r];
            sval[lp] = vv[r];
        }
    }
    __syncthreads();
    const int cnt = min(RS_TILE, n - tile * RS_TILE);
    for (int j = tid; j < cnt; j += RS_TPB) {
        const uint64_t k = skey[j];
        const int dg = (int)((k >> shift) & 255);
        const int pos = binpos[dg] + (j - tstart[dg]);
        kout[pos] = k;
        vout[pos] = sval[j];
    }
}

// ======================================================= collision runs
// Groups of sorted bodies whose keys are equal (all CSG_KEY_LEVELS levels):
// each member gets its levels CSG_KEY_LEVELS..63, then the members are re-ranked
// by (those levels, body index); ties keep index order.  Block j scans slice j
// of the sorted keys in chunks of CSG_TPB for the first members of runs (a run
// may extend past the slice; it belongs to the slice of its first member).  The
// thread that finds a run of at most CSG_RUN_SMALL members fixes it alone
// (fix_run_small); longer runs are fixed by the whole block.  Launched with one
// block per CSG_TPB bodies, so a slice is one chunk and blocks without a long run
// end after one look at their keys.
#define CSG_RUN_SMALL 16

// One thread's fix of the run [b, e) of sorted bodies (the thread owns the range).
__device__ void fix_run_small(int b, int e, int *__restrict__ idx, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ z,
                              const double *root, uint64_t *__restrict__ ext,
                              int *__restrict__ tmp_idx, uint64_t *__restrict__ tmp_ext) {
    for (int p = b; p < e; ++p) {
        const int i = idx[p];
        path_ext(x[i], y[i], z[i], root, ext + 3 * (size_t)p);
    }
    bool differ = false;
    for (int p = b + 1; p < e; ++p) {
        const uint64_t *u = ext + 3 * (size_t)p, *f = ext + 3 * (size_t)b;
        differ |= (u[0] != f[0]) | (u[1] != f[1]) | (u[2] != f[2]);
    }
    if (!differ) return;   // sorted by body index already (the sort is stable)
    for (int p = b; p < e; ++p) {
        const uint64_t *u = ext + 3 * (size_t)p;
        const int ip = idx[p];
        int rank = 0;
        for (int t = b; t < e; ++t) {
            const uint64_t *f = ext + 3 * (size_t)t;
            bool less;
            if (f[0] != u[0]) less = f[0] < u[0];
            else if (f[1] != u[1]) less = f[1] < u[1];
            else if (f[2] != u[2]) less = f[2] < u[2];
            else less = idx[t] < ip;
            rank += less ? 1 : 0;
        }
        tmp_idx[b + rank] = ip;
        for (int k = 0; k < 3; ++k) tmp_ext[3 * (size_t)(b + rank) + k] = u[k];
    }
    for (int p = b; p < e; ++p) {
        idx[p] = tmp_idx[p];
        for (int k = 0; k < 3; ++k) ext[3 * (size_t)p + k] = tmp_ext[3 * (size_t)p + k];
    }
}

__global__ __launch_bounds__(CSG_TPB) void k_runs(
    const uint64_t *__restrict__ key, int *__restrict__ idx, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, BuildParams *__restrict__ bp,
    uint64_t *__restrict__ ext, int *__restrict__ tmp_idx, uint64_t *__restrict__ tmp_ext) {
    __shared__ int s_end, s_nstart;
    __shared__ int s_start[CSG_TPB];
    const int nl = bp->n_live;
    double root[4] = {bp->root[0], bp->root[1], bp->root[2], bp->root[3]};
    const int per = max(CSG_TPB, (nl + (int)gridDim.x - 1) / (int)gridDim.x);
    const int lo = blockIdx.x * per, hi = min(nl, lo + per);
    int found = 0;
    for (int c0 = lo; c0 < hi; c0 += CSG_TPB) {
        if (threadIdx.x == 0) s_nstart = 0;
        __syncthreads();
        const int b = c0 + threadIdx.x;
        if (b < hi && b + 1 < nl) {
            const uint64_t k = key[b];
            if (k == key[b + 1] && (b == 0 || key[b - 1] != k)) {
                int e = b + 2;   // the run's end, followed while it stays short
                while (e < nl && e - b <= CSG_RUN_SMALL && key[e] == k) ++e;
                if (e - b <= CSG_RUN_SMALL) {
                    fix_run_small(b, e, idx, x, y, z, root, ext, tmp_idx, tmp_ext);
                    atomicAdd(&bp->n_runs, 1);
                } else {
                    s_start[atomicAdd(&s_nstart, 1)] = b;
                }
            }
        }
        __syncthreads();
        const int ns = s_nstart;
        found += ns;
        for (int q = 0; q < ns; ++q) {
            const int start = s_start[q];
            const uint64_t k0 = key[start];
            if (threadIdx.x == 0) s_end = nl;
            __syncthreads();
            // the run's end: the first body after start with another key
            for (int e0 = start + 1;; e0 += CSG_TPB) {
                const int j = e0 + threadIdx.x;
                const bool stop = (j >= nl) || (j < nl && key[j] != k0);
                if (stop) atomicMin(&s_end, j);
                if (__syncthreads_or(stop)) break;
            }
            const int end = min(s_end, nl);
            for (int p = start + threadIdx.x; p < end; p += CSG_TPB) {
                const int i = idx[p];
                path_ext(x[i], y[i], z[i], root, ext + 3 * (size_t)p);
            }
            __syncthreads();
            int differ = 0;
            for (int p = start + 1 + threadIdx.x; p < end; p += CSG_TPB) {
                const uint64_t *e = ext + 3 * (size_t)p, *f = ext + 3 * (size_t)start;
                differ |= (e[0] != f[0]) | (e[1] != f[1]) | (e[2] != f[2]);
            }
            if (__syncthreads_or(differ)) {
                for (int p = start + threadIdx.x; p < end; p += CSG_TPB) {
                    const uint64_t *e = ext + 3 * (size_t)p;
                    const int ip = idx[p];
                    int rank = 0;
                    for (int t = start; t < end; ++t) {
                        const uint64_t *f = ext + 3 * (size_t)t;
                        bool less;
                        if (f[0] != e[0]) less = f[0] < e[0];
                        else if (f[1] != e[1]) less = f[1] < e[1];
                        else if (f[2] != e[2]) less = f[2] < e[2];
                        else less = idx[t] < ip;
                        rank += less ? 1 : 0;
                    }
                    tmp_idx[start + rank] = ip;
                    tmp_ext[3 * (size_t)(start + rank) + 0] = e[0];
                    tmp_ext[3 * (size_t)(start + rank) + 1] = e[1];
                    tmp_ext[3 * (size_t)(start + rank) + 2] = e[2];
                }
                __syncthreads();
                for (int p = start + threadIdx.x; p < end; p += CSG_TPB) {
                    idx[p] = tmp_idx[p];
                    ext[3 * (size_t)p + 0] = tmp_ext[3 * (size_t)p + 0];
                    ext[3 * (size_t)p + 1] = tmp_ext[3 * (size_t)p + 1];
                    ext[3 * (size_t)p + 2] = tmp_ext[3 * (size_t)p + 2];
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0 && found) atomicAdd(&bp->n_runs, found);
}

// ================================================================== Peano-Hilbert wave order
// The FP32 wave-mask walk takes the 64 bodies of each wavefront in the
// Peano-Hilbert order of the tree's cells instead of the sorted (Morton) order.
// The tree and its node order stay as they are; only which bodies share a
// wavefront changes.  Consecutive Hilbert cells are face neighbours, so a
// wavefront's bodies are more compact and the union of their traversals is
// shorter (9% fewer wavefront node steps on the 1M-body merger at theta 0.5).
//
// c_hilbert[8 s + m] = (t << 3) | h: in a cell whose Hilbert orientation is
// state s, the child with octant digit m (bit 2 the first axis) comes h-th in
// Hilbert order and has state t.  State 0 is the root's.  The 24 states are those
// of Skilling's transform (J. Skilling, "Programming the Hilbert curve", AIP
// Conf. Proc. 707, 2004) applied to the octant digits as coordinate bits.
// The order is Hilbert down to CSG_KEY_LEVELS levels and sorted order inside a
// cell of that depth.  k_lcp writes each sorted body's Hilbert key, k_hilbert
// turns the tree's cells into rank offsets and k_hrank into the walk order.
static __constant__ unsigned char c_hilbert[24 * 8] = {
      8,  17,  27,   2,  39,  46,  52,   5,   // state 0
     56,  71,  73,  86,  91,  20,  10,  13,   // state 1
     48,   1, 103, 110, 115,  18,  12,  21,   // state 2
    126, 129,  29,  26,  79,  80, 140,   3,   // state 3
    148,  43,  37,  34, 127, 128,  78,  81,   // state 4
    156,  45,  35,  42,  31,   6, 160, 105,   // state 5
     72,  87, 139,   4,  57,  70,  50,  53,   // state 6
      0, 171, 111,  76,  49,  58, 102,  61,   // state 7
    180, 143,  83, 184,  69,  54,  66,  97,   // state 8
     16, 123,   9,  74,  47,  60,  38,  77,   // state 9
    132,  95,  85,  14,  67, 144,  82,  33,   // state 10
    142,  55, 185,  96,  93, 116,  90,  11,   // state 11
    188, 107, 175, 176, 101,  98,  62,  65,   // state 12
    164, 109, 119,  22,  99, 106, 152,  41,   // state 13
    174, 177,  63,  64, 117, 114,  92,  19,   // state 14
     30, 125, 161, 122,   7, 172, 104,  75,   // state 15
    130,  25, 133, 166, 179, 136,  84, 191,   // state 16
     94,  15, 141,  28, 145,  32, 138,  51,   // state 17
    146, 155, 149,  36, 137,  24, 190, 167,   // state 18
    154, 157, 147,  44, 169, 182, 120, 135,   // state 19
    162, 165, 121, 134, 187, 108, 168, 183,   // state 20
    118, 173,  23, 124, 153, 170,  40,  59,   // state 21
    178, 113, 131,  88, 181, 158,  68, 151,   // state 22
    186, 163,  89, 112, 189, 100, 150, 159,   // state 23
};

// Hilbert key of an octree key: its CSG_KEY_LEVELS octant digits walked through
// the state table (tab: c_hilbert staged in LDS), the first level in the top digit.
__device__ __forceinline__ uint64_t hilbert_key(uint64_t key, const unsigned char *tab) {
    uint64_t h = 0;
    int st = 0;
#pragma unroll
    for (int l = 0; l < CSG_KEY_LEVELS; ++l) {
        const int e = tab[8 * st + (int)((key >> (3 * (CSG_KEY_LEVELS - 1 - l))) & 7)];
        h = (h << 3) | (uint64_t)(e & 7);
        st = e >> 3;
    }
    return h;
}

// ============================================================ lcp + count
// L[b] = levels shared by sorted bodies b and b+1 (-1 past the end).  Body b
// opens the internal nodes of depths (min(L[b-1],63), min(L[b],63)] and starts
// a leaf unless it shares all 64 levels with b-1.
// Also M1[c] = min of L over the 64 bodies of chunk c (127 past the live
// bodies): the summary the node emission searches for a cell's last member;
// and hkey[b], the Peano-Hilbert key of sorted body b.
__global__ __launch_bounds__(CSG_TPB) void k_lcp(const uint64_t *__restrict__ key,
                                                 const uint64_t *__restrict__ ext,
                                                 BuildParams *__restrict__ bp,
                                                 signed char *__restrict__ L,
                                                 int *__restrict__ cnt,
                                                 int *__restrict__ part,
                                                 signed char *__restrict__ M1,
                                                 uint64_t *__restrict__ hkey) {
    __shared__ int wsum[4];
    __shared__ unsigned char tab[24 * 8];
    for (int k = threadIdx.x; k < 24 * 8; k += CSG_TPB) tab[k] = c_hilbert[k];
    __syncthreads();
    const int b = blockIdx.x * CSG_TPB + threadIdx.x;
    const int nl = bp->n_live;
    int lmin = 127, c = 0;
    if (b < nl) {
        hkey[b] = hilbert_key(key[b], tab);
        const int lc = (b + 1 < nl) ? lcp_sorted(key, ext, b, b + 1) : -1;
        const int lp = (b > 0) ? lcp_sorted(key, ext, b - 1, b) : -1;
        L[b] = (signed char)lc;
        lmin = lc;
        const int opened = max(0, min(lc, 63) - min(lp, 63));
        c = opened + (lp < 64 ? 1 : 0);
        cnt[b] = c;
    } else if (b == nl) {
        cnt[b] = 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lmin = min(lmin, __shfl_xor(lmin, o, 64));
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        M1[b >> 6] = (signed char)lmin;
        wsum[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    // the block's node count into its scan tile's counter (cleared by k_keys)
    if (threadIdx.x == 0)
        atomicAdd(&part[blockIdx.x / (CSG_TILE / CSG_TPB)], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

// Deepest leaf of the last tree, for the statistics only (bp->max_depth, cleared
// by k_bbox_final): a sorted body's leaf sits one level below the deeper of its
// common prefixes with its neighbours (64 at most).  CSG_TILE bodies per block.
__global__ __launch_bounds__(CSG_TPB) void k_max_depth(const signed char *__restrict__ L,
                                                       BuildParams *__restrict__ bp) {
    __shared__ int wmax[4];
    const int nl = bp->n_live;
    int d = 0;
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = blockIdx.x * CSG_TILE + r * CSG_TPB + threadIdx.x;
        if (b < nl) d = max(d, min(max(b > 0 ? (int)L[b - 1] : -1, (int)L[b]) + 1, 64));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d = max(d, __shfl_xor(d, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(&bp->max_depth, max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}

// Sorted copies of position and mass (coalesced reads for the scans and emit).
__global__ __launch_bounds__(CSG_TPB) void k_gather_sorted(
    const int *__restrict__ idx, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, const BuildParams *__restrict__ bp,
    double *__restrict__ xs, double *__restrict__ ys, double *__restrict__ zs,
    double *__restrict__ ms) {
    const int b = blockIdx.x * CSG_TPB + threadIdx.x;
    if (b >= bp->n_live) return;
    const int i = idx[b];
    xs[b] = x[i]; ys[b] = y[i]; zs[b] = z[i]; ms[b] = m[i];
}

// The same from the interleaved copy written by k_keys.
__global__ __launch_bounds__(CSG_TPB) void k_gather_sorted4(
    const int *__restrict__ idx, const double4 *__restrict__ xyzm,
    const BuildParams *__restrict__ bp, double *__restrict__ xs, double *__restrict__ ys,
    double *__restrict__ zs, double *__restrict__ ms) {
    const int b = blockIdx.x * CSG_TPB + threadIdx.x;
    if (b >= bp->n_live) return;
    const double4 v = xyzm[idx[b]];
    xs[b] = v.x; ys[b] = v.y; zs[b] = v.z; ms[b] = v.w;
}

// ============================================================ int scan
// Per-tile sums of in[0..n_live], CSG_TILE entries per block, into part.
__global__ __launch_bounds__(CSG_TPB) void k_iscan_reduce(const int *__restrict__ in,
                                                          const BuildParams *__restrict__ bp,
                                                          int *__restrict__ part) {
    __shared__ int lds4[4];
    const int len = bp->n_live + 1;
    const int base = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int s = 0;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) s += (base + r < len) ? in[base + r] : 0;
    int total;
    (void)block_excl_scan(s, lds4, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Exclusive scan of in[0..n_live] (the node counts) into out, one tile of CSG_TILE
// per block: the tile sums k_lcp accumulated in part (the tiles before this one,
// summed here) and a scan inside the tile; the entry past the last body is the
// node count.  Block t also writes M2[t], the minimum of the 64 chunk minima
// M1[64 t, 64 t + 64) (the superchunk summary cell_last searches).
__global__ __launch_bounds__(CSG_TPB) void k_iscan_down(const int *__restrict__ in,
                                                        const int *__restrict__ part,
                                                        BuildParams *bp, int *__restrict__ out,
                                                        const signed char *__restrict__ M1,
                                                        int nm1, signed char *__restrict__ M2) {
    __shared__ int lds4[4];
    if (threadIdx.x < 64 && (int)blockIdx.x * 64 < nm1) {
        const int i = blockIdx.x * 64 + threadIdx.x;
        int m = i < nm1 ? (int)M1[i] : 127;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
        if (threadIdx.x == 0) M2[blockIdx.x] = (signed char)m;
    }
    const int nl = bp->n_live, len = nl + 1;
    int c = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += CSG_TPB) c += part[k];
    int carry;
    (void)block_excl_scan(c, lds4, carry);
    const int base = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int v[CSG_ITEMS], s = 0;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) { v[r] = (base + r < len) ? in[base + r] : 0; s += v[r]; }
    int total;
    int run = block_excl_scan(s, lds4, total) + carry;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int i = base + r;
        if (i < len) {
            out[i] = run;
            if (i == nl) {
                const int nn = (nl == 0) ? 1 : run;
                bp->n_nodes = nn;
                if (nn + CSG_NODE_PAD > bp->node_cap) bp->overflow = 1;
            }
        }
        run += v[r];
    }
}

// ====================================================== double-double scan
// Exclusive prefix P[b] over sorted bodies of (m, m*x, m*y, m*z), each a DD;
// P[n_live] is the total.  Layout: pref[b*4 + c], c = 0..3 (one 64-byte line per b).  Every sum is
// formed in a fixed order (thread-serial over 8 elements, then a wave shuffle
// scan, then the 4 waves in order), so the result is reproducible.
__device__ __forceinline__ void dd_elem(const double *__restrict__ xs,
                                        const double *__restrict__ ys,
                                        const double *__restrict__ zs,
                                        const double *__restrict__ ms, int b, int nl, DD v[4]) {
    if (b < nl) {
        const double mm = ms[b];
        v[0] = DD{mm, 0.0};
        v[1] = dd_prod(mm, xs[b]);
        v[2] = dd_prod(mm, ys[b]);
        v[3] = dd_prod(mm, zs[b]);
    } else {
        for (int c = 0; c < 4; ++c) v[c] = DD{0.0, 0.0};
    }
}

// Block-wide inclusive DD scan of one value per thread; `wtot` is LDS for the
// four wave totals.
__device__ __forceinline__ void block_dd_incl_scan(DD v[4], DD (*wtot)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const DD t{__shfl_up(v[c].h, o, 64), __shfl_up(v[c].l, o, 64)};
            if (lane >= o) v[c] = dd_add(t, v[c]);
        }
    }
    if (lane == 63)
        for (int c = 0; c < 4; ++c) wtot[w][c] = v[c];
    __syncthreads();
    if (w > 0) {
        for (int c = 0; c < 4; ++c) {
            DD off = wtot[0][c];
            for (int k = 1; k < w; ++k) off = dd_add(off, wtot[k][c]);
            v[c] = dd_add(off, v[c]);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void dd_thread_sum(const double *__restrict__ xs,
                                              const double *__restrict__ ys,
                                              const double *__restrict__ zs,
                                              const double *__restrict__ ms, int base, int nl,
                                              DD s[4]) {
    for (int c = 0; c < 4; ++c) s[c] = DD{0.0, 0.0};
    for (int r = 0; r < CSG_ITEMS; ++r) {
        DD e[4];
        dd_elem(xs, ys, zs, ms, base + r, nl, e);
        for (int c = 0; c < 4; ++c) s[c] = dd_add(s[c], e[c]);
    }
}

__global__ __launch_bounds__(CSG_TPB) void k_dd_reduce(const double *__restrict__ xs,
                                                       const double *__restrict__ ys,
                                                       const double *__restrict__ zs,
                                                       const double *__restrict__ ms,
                                                       const BuildParams *__restrict__ bp,
                                                       DD *__restrict__ part) {
    __shared__ DD wtot[4][4];
    const int nl = bp->n_live;
    DD s[4];
    dd_thread_sum(xs, ys, zs, ms, blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS, nl, s);
    block_dd_incl_scan(s, wtot);
    if (threadIdx.x == CSG_TPB - 1)
        for (int c = 0; c < 4; ++c) part[blockIdx.x * 4 + c] = s[c];
}

// Exclusive scan of the per-block DD totals (one block, chunks of 256).
__global__ __launch_bounds__(CSG_TPB) void k_dd_top(DD *__restrict__ part, int nb) {
    __shared__ DD wtot[4][4];
    __shared__ DD prev[CSG_TPB][4];
    DD carry[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int c0 = 0; c0 < nb; c0 += CSG_TPB) {
        const int i = c0 + threadIdx.x;
        DD v[4];
        for (int c = 0; c < 4; ++c) v[c] = (i < nb) ? part[i * 4 + c] : DD{0, 0};
        block_dd_incl_scan(v, wtot);
        for (int c = 0; c < 4; ++c) prev[threadIdx.x][c] = v[c];
        __syncthreads();
        if (i < nb)
            for (int c = 0; c < 4; ++c)
                part[i * 4 + c] =
                    threadIdx.x > 0 ? dd_add(carry[c], prev[threadIdx.x - 1][c]) : carry[c];
        for (int c = 0; c < 4; ++c) carry[c] = dd_add(carry[c], prev[CSG_TPB - 1][c]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(CSG_TPB) void k_dd_down(const double *__restrict__ xs,
                                                     const double *__restrict__ ys,
                                                     const double *__restrict__ zs,
                                                     const double *__restrict__ ms,
                                                     const BuildParams *__restrict__ bp,
                                                     const DD *__restrict__ part,
                                                     DD *__restrict__ pref, int cap1) {
    __shared__ DD wtot[4][4];
    __shared__ DD prev[CSG_TPB][4];
    const int nl = bp->n_live;
    const int base = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    DD s[4];
    dd_thread_sum(xs, ys, zs, ms, base, nl, s);
    block_dd_incl_scan(s, wtot);
    for (int c = 0; c < 4; ++c) prev[threadIdx.x][c] = s[c];
    __syncthreads();
    DD run[4];
    for (int c = 0; c < 4; ++c)
        run[c] = threadIdx.x > 0 ? dd_add(part[blockIdx.x * 4 + c], prev[threadIdx.x - 1][c])
                                 : part[blockIdx.x * 4 + c];
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = base + r;
        if (b <= nl)
            for (int c = 0; c < 4; ++c) pref[(size_t)b * 4 + c] = run[c];
        if (b < nl) {
            DD e[4];
            dd_elem(xs, ys, zs, ms, b, nl, e);
            for (int c = 0; c < 4; ++c) run[c] = dd_add(run[c], e[c]);
        }
    }
}

// ------------------------------------------- fused gather + tile sums, down
// The build's form of the three kernels above.  k_gather_dd writes the sorted
// copies of (x, y, z, m) from the interleaved body-order copy and forms, from
// the same registers, the double-double sums of each thread's 8 consecutive
// sorted bodies, each thread's exclusive prefix inside its tile (tpref) and the
// tile total (part).  k_dd_down2 then needs no scan of its own: each block adds
// the totals of the tiles before it (a fixed-order reduction, so the result is
// reproducible) and runs the per-body prefixes from tpref.  Global accesses go
// through LDS so that consecutive lanes touch consecutive bodies; the LDS index
// is padded by one slot per 8 elements, so the per-thread accesses (stride 8)
// fall in distinct banks.
#define DD_PAD(j) ((j) + ((j) >> 3))
#define DD_TILE_LDS (CSG_TILE + CSG_TILE / 8)

__device__ __forceinline__ void dd_elem4(double4 q, bool live, DD v[4]) {
    if (live) {
        v[0] = DD{q.w, 0.0};
        v[1] = dd_prod(q.w, q.x);
        v[2] = dd_prod(q.w, q.y);
        v[3] = dd_prod(q.w, q.z);
    } else {
        for (int c = 0; c < 4; ++c) v[c] = DD{0.0, 0.0};
    }
}

__global__ __launch_bounds__(CSG_TPB) void k_gather_dd(
    const int *__restrict__ idx, const double4 *__restrict__ xyzm,
    const BuildParams *__restrict__ bp, double *__restrict__ xs, double *__restrict__ ys,
    double *__restrict__ zs, double *__restrict__ ms, DD *__restrict__ part,
    DD *__restrict__ tpref) {
    __shared__ double sbuf[DD_TILE_LDS];
    __shared__ int sidx[DD_TILE_LDS];
    __shared__ DD wtot[4][4];
    __shared__ DD prev[CSG_TPB][4];
    const int nl = bp->n_live;
    const int t = threadIdx.x;
    const int tile0 = blockIdx.x * CSG_TILE;
#pragma unroll
    for (int k = 0; k < CSG_ITEMS; ++k) {
        const int j = t + k * CSG_TPB;
        sidx[DD_PAD(j)] = tile0 + j < nl ? idx[tile0 + j] : 0;
    }
    __syncthreads();
    const int e0 = tile0 + t * CSG_ITEMS;
    double4 q[CSG_ITEMS];
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r)
        q[r] = e0 + r < nl ? xyzm[sidx[DD_PAD(t * CSG_ITEMS + r)]]
                           : make_double4(0.0, 0.0, 0.0, 0.0);
    DD s[4];
    for (int c = 0; c < 4; ++c) s[c] = DD{0.0, 0.0};
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        DD e[4];
        dd_elem4(q[r], e0 + r < nl, e);
        for (int c = 0; c < 4; ++c) s[c] = dd_add(s[c], e[c]);
    }
    // sorted copies, one coordinate at a time through LDS
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int r = 0; r < CSG_ITEMS; ++r)
            sbuf[DD_PAD(t * CSG_ITEMS + r)] = a == 0 ? q[r].x : a == 1 ? q[r].y : a == 2 ? q[r].z : q[r].w;
        __syncthreads();
        double *const dst = a == 0 ? xs : a == 1 ? ys : a == 2 ? zs : ms;
#pragma unroll
        for (int k = 0; k < CSG_ITEMS; ++k) {
            const int j = t + k * CSG_TPB;
            if (tile0 + j < nl) dst[tile0 + j] = sbuf[DD_PAD(j)];
        }
        __syncthreads();
    }
    block_dd_incl_scan(s, wtot);
    for (int c = 0; c < 4; ++c) prev[t][c] = s[c];
    __syncthreads();
    DD *const tp = tpref + ((size_t)blockIdx.x * CSG_TPB + t) * 4;
    for (int c = 0; c < 4; ++c) tp[c] = t > 0 ? prev[t - 1][c] : DD{0.0, 0.0};
    if (t == CSG_TPB - 1)
        for (int c = 0; c < 4; ++c) part[blockIdx.x * 4 + c] = s[c];
}

// Also clears the Peano-Hilbert difference array hdiff[0..n_live] that k_emit then
// accumulates into (the node-count scan's buffer, which it is done with by now).
__global__ __launch_bounds__(CSG_TPB) void k_dd_down2(
    const double *__restrict__ xs, const double *__restrict__ ys, const double *__restrict__ zs,
    const double *__restrict__ ms, const BuildParams *__restrict__ bp,
    const DD *__restrict__ part, const DD *__restrict__ tpref, DD *__restrict__ pref,
    int *__restrict__ hdiff) {
    __shared__ double sbuf[DD_TILE_LDS];
    __shared__ DD red[CSG_TPB][4];
    const int nl = bp->n_live;
    const int t = threadIdx.x;
    const int tile0 = blockIdx.x * CSG_TILE;
    // every global load first (one round trip): this thread's tile prefix, the
    // coordinates (coalesced), the totals of the tiles before this one
    const DD *const tp = tpref + ((size_t)blockIdx.x * CSG_TPB + t) * 4;
    DD run[4];
    for (int c = 0; c < 4; ++c) run[c] = tp[c];
    double g[4][CSG_ITEMS];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double *const src = a == 0 ? xs : a == 1 ? ys : a == 2 ? zs : ms;
#pragma unroll
        for (int k = 0; k < CSG_ITEMS; ++k) {
            const int j = t + k * CSG_TPB;
            g[a][k] = tile0 + j < nl ? src[tile0 + j] : 0.0;
        }
    }
    DD acc[4];
    for (int c = 0; c < 4; ++c) acc[c] = DD{0.0, 0.0};
    for (int k = t; k < (int)blockIdx.x; k += CSG_TPB)
        for (int c = 0; c < 4; ++c) acc[c] = dd_add(acc[c], part[k * 4 + c]);
    for (int c = 0; c < 4; ++c) red[t][c] = acc[c];
    // this thread's 8 consecutive bodies, one coordinate at a time through LDS
    double v[4][CSG_ITEMS];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int k = 0; k < CSG_ITEMS; ++k) sbuf[DD_PAD(t + k * CSG_TPB)] = g[a][k];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < CSG_ITEMS; ++r) v[a][r] = sbuf[DD_PAD(t * CSG_ITEMS + r)];
        __syncthreads();
    }
    // tile offset: fixed-order tree over the strided partial sums
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if (t < st)
            for (int c = 0; c < 4; ++c) red[t][c] = dd_add(red[t][c], red[t + st][c]);
        __syncthreads();
    }
    for (int c = 0; c < 4; ++c) run[c] = dd_add(red[0][c], run[c]);
    const int e0 = tile0 + t * CSG_ITEMS;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = e0 + r;
        if (b <= nl) {
            for (int c = 0; c < 4; ++c) pref[(size_t)b * 4 + c] = run[c];
            hdiff[b] = 0;
        }
        if (b < nl) {
            DD e[4];
            dd_elem4(make_double4(v[0][r], v[1][r], v[2][r], v[3][r]), true, e);
            for (int c = 0; c < 4; ++c) run[c] = dd_add(run[c], e[c]);
        }
    }
}

__device__ __forceinline__ double dd_range(const DD *__restrict__ pref, int cap1, int c, int b,
                                           int e) {
    (void)cap1;
    const DD hi = pref[(size_t)(e + 1) * 4 + c], lo = pref[(size_t)b * 4 + c];
    return dd_add(hi, dd_neg(lo)).h;
}

// ================================================================== emit
// First i in [from, to) with a[i] < d, else to.  Reads the whole 16-byte words
// that overlap [from, to): the arrays are 16-byte aligned and padded to 16 bytes.
__device__ __forceinline__ int first_below(const signed char *__restrict__ a, int from, int to,
                                           int d) {
    for (int w = from & ~15; w < to; w += 16) {
        const int4 q = *(const int4 *)(a + w);
        const int vv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int i = w + j;
            const int v = (signed char)(vv[j >> 2] >> (8 * (j & 3)));
            if (i >= from && i < to && v < d) return i;
        }
    }
    return to;
}

// Last member of the cell of depth d whose first member is sorted body b: the
// first k >= b whose common prefix with k + 1 is shallower than d (L[k] < d;
// L[nl - 1] = -1 ends every search).  Three levels: the rest of b's chunk of
// 64 bodies in L, the rest of its superchunk of 4096 in the chunk minima M1,
// then the superchunk minima M2, descending again to the chunk and the body.
__device__ int cell_last(const signed char *__restrict__ L, const signed char *__restrict__ M1,
                         const signed char *__restrict__ M2, int b, int d, int nl) {
    const int cend = min((b | 63) + 1, nl);
    const int k = first_below(L, b, cend, d);
    if (k < cend) return k;
    const int nch = (nl + 63) >> 6;
    const int sc = b >> 12;
    const int send = min((sc + 1) << 6, nch);
    int c = first_below(M1, (b >> 6) + 1, send, d);
    if (c == send) {
        const int s2 = first_below(M2, sc + 1, (nch + 63) >> 6, d);
        c = first_below(M1, s2 << 6, min((s2 + 1) << 6, nch), d);
    }
    return first_below(L, c << 6, min((c + 1) << 6, nl), d);
}

// The internal nodes the 64 bodies of a wavefront open, dealt over its lanes: lane
// j takes the wave's j-th node (in body order, then depth) and calls f(b, d, id)
// with the node's first member, depth and id; opened / lo / id0 are this lane's
// body's node count, first depth and first id.  A body that opens many nested
// cells, like body 0 with the ancestors of the first leaf, does not walk them
// one after another.  Every thread of the block calls it (one barrier).
template <class F>
__device__ __forceinline__ void deal_wave(int opened, int lo, int id0, int *s_off, int *s_lo,
                                          int *s_id, F &&f) {
    const int lane = threadIdx.x & 63, wbase = threadIdx.x & ~63;
    int incl = opened;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    const int total = __shfl(incl, 63, 64);
    s_off[threadIdx.x] = incl - opened;
    s_lo[threadIdx.x] = lo;
    s_id[threadIdx.x] = id0;
    __syncthreads();
    for (int j = lane; j < total; j += 64) {
        // the last lane whose offset is <= j opens node j (it opens at least one)
        int k0 = 0, k1 = 63;
        while (k0 < k1) {
            const int mid = (k0 + k1 + 1) >> 1;
            if (s_off[wbase + mid] <= j) k0 = mid; else k1 = mid - 1;
        }
        const int r = j - s_off[wbase + k0];
        f((int)(blockIdx.x * CSG_TPB) + wbase + k0, s_lo[wbase + k0] + r, s_id[wbase + k0] + r);
    }
}

// Node records.  The nodes a sorted body b opens are depths (min(L[b-1],63),
// min(L[b],63)] (internal cells whose first member is b) plus its leaf, with
// consecutive pre-order ids from nbase[b].  k_emit writes the leaves and, for
// k_emit writes both: each body its leaf, and the internal nodes that the 64
// bodies of a wavefront open are dealt out over the wave's lanes (lane j takes
// the wave's j-th one), so a body that opens many nested cells, like body 0 with
// the ancestors of the first leaf, does not walk them one after another.  Each
// internal node finds its last member with cell_last.
__device__ __forceinline__ void emit_internal(
    const signed char *__restrict__ L, const signed char *__restrict__ M1,
    const signed char *__restrict__ M2, const int *__restrict__ nbase,
    const DD *__restrict__ pref, int cap1, const BuildParams *__restrict__ bp, int nl, int b,
    int d, int id, WNode *__restrict__ nodes, WNode32 *__restrict__ n32, double inv_th2,
    int *__restrict__ nlast) {
    const double H = bp->root[3];
    const int e = (d == 0) ? nl - 1 : cell_last(L, M1, M2, b, d, nl);
    nlast[id] = e;
    WNode w;
    const double M = dd_range(pref, cap1, 0, b, e);
    w.cx = dd_range(pref, cap1, 1, b, e) / M;
    w.cy = dd_range(pref, cap1, 2, b, e) / M;
    w.cz = dd_range(pref, cap1, 3, b, e) / M;
    w.m = M;
    const double s = 2.0 * ldexp(H, -d);
    w.s2 = s * s;
    w.skip = nbase[e + 1];
    w.body = -1; w.depth = d; w.first = b; w.last = e; w.pad = 0;
    store_nodes(nodes, n32, id, w, inv_th2);
}

__global__ __launch_bounds__(CSG_TPB) void k_emit(
    const uint64_t *__restrict__ key, const uint64_t *__restrict__ ext,
    const int *__restrict__ idx, const signed char *__restrict__ L,
    const signed char *__restrict__ M1, const signed char *__restrict__ M2,
    const int *__restrict__ nbase, const DD *__restrict__ pref, int cap1,
    const double *__restrict__ xs, const double *__restrict__ ys, const double *__restrict__ zs,
    const double *__restrict__ ms, const BuildParams *__restrict__ bp, WNode *__restrict__ nodes,
    WNode32 *__restrict__ n32, double inv_th2, int *__restrict__ nlast) {
    __shared__ int s_off[CSG_TPB], s_lo[CSG_TPB], s_id[CSG_TPB];
    const int b = blockIdx.x * CSG_TPB + threadIdx.x;
    const int nl = bp->n_live;
    if (bp->overflow) return;
    const double H = bp->root[3];
    if (b == 0) {
        // padding records after the tree (the FP32 walk's unrolled rounds may
        // step onto up to seven of them): opening criterion never met
        // (s2 = +inf), so no lane interacts, descends past them one by one
        const int nn = bp->n_nodes;
        for (int k = 0; k < CSG_NODE_PAD; ++k) {
            WNode32 f;
            f.cx = f.cy = f.cz = 0.0;
            f.m = 0.0f;
            f.s2 = __builtin_inff();
            f.skip = nn + k + 1;
            f.body = -1;
            n32[nn + k] = f;
            if (nodes) {   // the same for the FP64 wave-mask walk (no mass: never interacts)
                WNode w;
                w.cx = w.cy = w.cz = 0.0;
                w.m = 0.0;
                w.s2 = __builtin_inf();
                w.skip = nn + k + 1;
                w.body = -1; w.depth = 0; w.first = 0; w.last = -1; w.pad = 0;
                nodes[nn + k] = w;
            }
        }
    }
    if (nl == 0) {
        if (b == 0) {
            WNode r;
            r.cx = r.cy = r.cz = 0.0;
            r.m = 0.0;
            const double s = 2.0 * H;
            r.s2 = s * s;
            r.skip = 1; r.body = -1; r.depth = 0; r.first = 0; r.last = -1; r.pad = 0;
            store_nodes(nodes, n32, 0, r, inv_th2);
        }
        return;
    }
    const bool live = b < nl;
    int lp = -1, lc = -1, lo = 0, id0 = 0, opened = 0;
    if (live) {
        lp = (b > 0) ? L[b - 1] : -1;
        lc = L[b];
        lo = min(lp, 63) + 1;
        opened = max(0, min(lc, 63) - lo + 1);   // internal depths lo .. min(lc, 63)
        id0 = nbase[b];
    }
    deal_wave(opened, lo, id0, s_off, s_lo, s_id, [&](int bb, int d, int id) {
        emit_internal(L, M1, M2, nbase, pref, cap1, bp, nl, bb, d, id, nodes, n32, inv_th2, nlast);
    });
    if (live && lp < 64) {
        const int id = id0 + opened;
        WNode w;
        const int i = idx[b];
        int e = b;
        int depth;
        if (lc >= 64) {
            depth = 64;
            e = range_end(key, ext, b, 64, nl);
            const double M = dd_range(pref, cap1, 0, b, e);
            w.cx = dd_range(pref, cap1, 1, b, e) / M;
            w.cy = dd_range(pref, cap1, 2, b, e) / M;
            w.cz = dd_range(pref, cap1, 3, b, e) / M;
            w.m = M;
        } else {
            depth = max(lp, lc) + 1;
            w.cx = xs[b]; w.cy = ys[b]; w.cz = zs[b];
            w.m = ms[b];
        }
        const double s = 2.0 * ldexp(H, -depth);
        w.s2 = s * s;
        w.skip = id + 1;
        w.body = i; w.depth = depth; w.first = b; w.last = e; w.pad = 0;
        store_nodes(nodes, n32, id, w, inv_th2);
        nlast[id] = e;
    }
}

// Peano-Hilbert offsets of the cells of depth < CSG_KEY_LEVELS (dealt like
// k_emit's nodes): a cell's children are found through the last members k_emit
// recorded (the first child is the next node in pre-order, a later one the first
// node its first member opens), their Hilbert digits from the members' Hilbert
// keys.  Each child's Hilbert offset within the cell minus its sorted offset is
// added over its members as a difference array: the change from one child's
// value to the next at each boundary between two children, stored in hbnd (a
// position k is such a boundary of exactly one cell, the one of depth L[k-1]),
// the first child's value at the cell's first member and minus the last child's
// past its last member, added into hdiff (cells share those positions).  A
// sorted body's Hilbert rank is its position plus the prefix sum up to it
// (hilbert_step).
__global__ __launch_bounds__(CSG_TPB) void k_hilbert(const signed char *__restrict__ L,
                                                     const int *__restrict__ nbase,
                                                     const int *__restrict__ nlast,
                                                     const uint64_t *__restrict__ hkey,
                                                     const BuildParams *__restrict__ bp,
                                                     int *__restrict__ hdiff,
                                                     int *__restrict__ hbnd) {
    __shared__ int s_off[CSG_TPB], s_lo[CSG_TPB], s_id[CSG_TPB];
    constexpr int KL = CSG_KEY_LEVELS;
    const int b = blockIdx.x * CSG_TPB + threadIdx.x;
    const int nl = bp->n_live;
    if (bp->overflow || nl == 0) return;
    int lo = 0, id0 = 0, opened = 0;
    if (b < nl) {
        const int lp = (b > 0) ? L[b - 1] : -1;
        lo = min(lp, 63) + 1;
        opened = max(0, min((int)L[b], KL - 1) - lo + 1);   // depths lo .. min(L[b], KL - 1)
        id0 = nbase[b];
    }
    deal_wave(opened, lo, id0, s_off, s_lo, s_id, [&](int s, int d, int id) {
        const int e = nlast[id];
        int f[9], h[8];
        int k0 = s;
        bool done = false;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (done) { f[j] = e + 1; h[j] = 8; continue; }
            f[j] = k0;
            h[j] = (int)((hkey[k0] >> (3 * (KL - 1 - d))) & 7);
            const int ce = nlast[j == 0 ? id + 1 : nbase[k0]];
            if (ce >= e || ce < k0) done = true; else k0 = ce + 1;   // k0 stays in [s, e]
        }
        f[8] = e + 1;
        if (f[1] > e) return;   // one child: it keeps its place
        int prev = 0;
#pragma unroll
        for (int j = 0; j <= 8; ++j) {
            const bool in = j < 8 && f[j] <= e;
            int cur = 0;
            if (in) {
                int off = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (h[i] < h[j]) off += f[i + 1] - f[i];
                cur = off - (f[j] - s);
            }
            // f[j] <= e + 1 <= n_live
            if (j > 0 && in) hbnd[f[j]] = cur - prev;
            else if (cur != prev) atomicAdd(&hdiff[f[j]], cur - prev);
            prev = cur;
            if (!in) break;
        }
    });
}

// The rank step at sorted position k < n_live: hdiff[k], plus hbnd[k] when k is a
// boundary between two children of a cell of depth < CSG_KEY_LEVELS.
__device__ __forceinline__ int hilbert_step(const int *__restrict__ hdiff,
                                            const int *__restrict__ hbnd,
                                            const signed char *__restrict__ L, int k) {
    int v = hdiff[k];
    if (k > 0) {
        const int d = L[k - 1];
        if (d >= 0 && d < CSG_KEY_LEVELS) v += hbnd[k];
    }
    return v;
}

// Per-tile sums of the rank steps of positions < n_live, CSG_TILE per block.
__global__ __launch_bounds__(CSG_TPB) void k_hsum(const int *__restrict__ hdiff,
                                                  const int *__restrict__ hbnd,
                                                  const signed char *__restrict__ L,
                                                  const BuildParams *__restrict__ bp,
                                                  int *__restrict__ part) {
    __shared__ int lds4[4];
    const int nl = bp->n_live;
    const int base = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int s = 0;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r)
        if (base + r < nl) s += hilbert_step(hdiff, hbnd, L, base + r);
    int total;
    (void)block_excl_scan(s, lds4, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Peano-Hilbert rank of every sorted body b < n_live: b plus the prefix sum of
// its rank steps (the tile sums hpart of k_hsum before this tile, then a scan
// inside it); the walk order gets widx[rank] = idx[b] and, when hpos is given, the
// sorted position hpos[rank] = b (how a tape recorded in this order maps its
// wavefronts' lanes back to sorted bodies).  A rank outside [0, n_live) (a broken
// difference array) is not stored and flags order_bad, which fails the build.
__global__ __launch_bounds__(CSG_TPB) void k_hrank(const int *__restrict__ hdiff,
                                                   const int *__restrict__ hbnd,
                                                   const signed char *__restrict__ L,
                                                   const int *__restrict__ hpart,
                                                   const int *__restrict__ idx,
                                                   BuildParams *__restrict__ bp,
                                                   int *__restrict__ widx,
                                                   int *__restrict__ hpos) {
    __shared__ int lds4[4];
    const int nl = bp->n_live;
    if (bp->overflow) return;
    const int tile0 = blockIdx.x * CSG_TILE;
    if (tile0 >= nl) return;
    int c = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += CSG_TPB) c += hpart[k];
    int carry;
    (void)block_excl_scan(c, lds4, carry);
    const int base = tile0 + threadIdx.x * CSG_ITEMS;
    int v[CSG_ITEMS], sum = 0;
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        v[r] = (base + r < nl) ? hilbert_step(hdiff, hbnd, L, base + r) : 0;
        sum += v[r];
    }
    int total;
    int run = carry + block_excl_scan(sum, lds4, total);
#pragma unroll
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = base + r;
        if (b < nl) {
            run += v[r];
            const int rank = b + run;
            if ((unsigned)rank < (unsigned)nl) {
                widx[rank] = idx[b];
                if (hpos) hpos[rank] = b;
            } else {
                bp->order_bad = 1;
            }
        }
    }
}

// ================================================================== walk
enum { WALK_FP64_STRICT = 0, WALK_FP64 = 1, WALK_FP32 = 2 };

// Reference opening test fl(s2 / d2) < theta2 (src/octree.c:202-203) decided
// from one fused multiply-add; the IEEE division runs only within 1e-14 of the
// boundary, where the sign of s2 - theta2*d2 could disagree with it.
__device__ __forceinline__ bool open_accept_exact(double s2, double d2, double theta2) {
    if (!(d2 > 0.0)) return false;
    const double r = fma(-theta2, d2, s2);
    const double tol = 1e-14 * s2;
    if (r < -tol) return true;
    if (r > tol) return false;
    return (s2 / d2) < theta2;
}

// Work-group -> tile mapping.  Work groups are dealt round-robin over the 8
// XCDs (each with its own L2): block b runs on XCD b % 8.  Consecutive tiles
// are kept on one XCD in chunks of `chunk` tiles, so neighbouring waves share
// that XCD's L2, while the chunks of one XCD are spread over the whole body
// range to balance dense and sparse regions.  chunk == 0: one contiguous range
// per XCD; chunk < 0: tile = block.  The grid is a multiple of 8 * chunk;
// surplus tiles exit.
__device__ __forceinline__ int xcd_tile(int grid, int chunk) {
    const int x = blockIdx.x & 7, q = blockIdx.x >> 3;
    if (chunk < 0) return blockIdx.x;
    if (chunk == 0) return x * (grid >> 3) + q;
    return (q / chunk) * (8 * chunk) + x * chunk + (q % chunk);
}

// Wave index of the calling wavefront in the walks: through the dispatch order
// when one is given (slot s of the grid -> order[s]), else the XCD-chunked tile
// mapping.  -1: a slot past the order's end (no wave).
__device__ __forceinline__ int walk_wave(int nb8, int chunk, const WaveOrder &wo) {
    if (wo.order) {
        const int slot = (int)blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
        if (slot >= wo.slots) return -1;
        const int w = wo.order[slot];
        return (w < 0 || w >= wo.n) ? -1 : w;
    }
    return xcd_tile(nb8, chunk) * (CSG_TPB / 64) + (int)(threadIdx.x >> 6);
}

// Per-lane state of the FP64 walks.  A lane holds one body; p is the next
// pre-order node that this body must look at (its own DFS position).
struct WalkLane {
    int me;              // original body index (-1 for an idle lane)
    double x, y, z;      // body position
    int p;
    int cnt;
    bool inter;          // whether the last visited node interacted
    double sx, sy, sz;   // accumulated acceleration (without G in the fast walk)
};

// Body-node step at node i (src/octree.c:182-217): skip an empty cell, a leaf
// interacts unless it is the body itself, an internal node is accepted
// (interacts) or opened.  Returns whether this lane opens the node.
template <int PREC>
__device__ __forceinline__ bool walk_visit(WalkLane &L, const WNode &nd, int i, double G,
                                           double eps2, double theta2) {
    bool desc = false;
    const double dx = nd.cx - L.x, dy = nd.cy - L.y, dz = nd.cz - L.z;
    if constexpr (PREC == WALK_FP64_STRICT) {
        // reference arithmetic and evaluation order, no contraction
        const double d2 = dx * dx + dy * dy + dz * dz;
        bool inter;
        if (nd.body >= 0) inter = nd.body != L.me;
        else { inter = open_accept_exact(nd.s2, d2, theta2); desc = !inter; }
        if (nd.m == 0.0) { inter = false; desc = false; }
        L.inter = inter;
        if (inter) {
            const double r2 = d2 + eps2;
            const double w = 1.0 / (r2 * sqrt(r2));
            const double gm = G * nd.m;
            L.sx += gm * dx * w;
            L.sy += gm * dy * w;
            L.sz += gm * dz * w;
            ++L.cnt;
        }
    } else {
        // refined hardware rsqrt, FMA, G applied once at the end; every cell of
        // a tree that has walkers has mass > 0
        const double d2 = fma(dx, dx, fma(dy, dy, dz * dz));
        bool inter;
        if (nd.body >= 0) inter = nd.body != L.me;
        else { inter = open_accept_exact(nd.s2, d2, theta2); desc = !inter; }
        L.inter = inter;
        if (inter) {
            const double r2 = d2 + eps2;
            const double yv = rsq64_walk(r2);
            const double f = nd.m * (yv * yv * yv);
            L.sx = fma(dx, f, L.sx);
            L.sy = fma(dy, f, L.sy);
            L.sz = fma(dz, f, L.sz);
            ++L.cnt;
        }
    }
    L.p = desc ? i + 1 : nd.skip;
    return desc;
}

// FP64 walks.  One wavefront per 64 consecutive sorted (Morton-adjacent)
// bodies, split into 64/GS traversal groups of GS lanes.  A group walks the
// pre-order node array with one shared index i: it visits node i when any of
// its lanes has p == i, and moves to i+1 when any lane opened the node, else to
// skip[i].  Lanes never visit a node outside their own DFS sequence, so each
// body's interactions and their summation order are the reference's.  GS = 64
// keeps i in a scalar register and fetches nodes with scalar loads; smaller
// groups keep several fetches in flight per wave.
template <int PREC, int GS>
__global__ __launch_bounds__(CSG_TPB) void k_walk(
    const WNode *__restrict__ nodes, const BuildParams *__restrict__ bp,
    const int *__restrict__ sidx, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    double G, double eps2, double theta2, int nb8, int chunk, TapeOut tape, WaveOrder wo) {
    const int n_live = bp->n_live;
    const int nnodes = bp->n_nodes;
    if (bp->overflow) return;
    const int lane = threadIdx.x & 63;
    const int wv = walk_wave(nb8, chunk, wo);   // wave w walks sorted bodies [64 w, 64 w + 64)
    if (wv < 0) return;
    const int b0 = wv * 64;
    if (b0 >= n_live) {
        if (wo.cost && lane == 0 && wv < wo.n) wo.cost[wv] = 0;
        return;
    }
    const int b = b0 + lane;
    const bool act = b < n_live;
    const bool taping = (GS == 64) && tape.rec != nullptr;
    int nrec = 0, nit = 0;
    WalkLane L;
    L.me = act ? sidx[b] : -1;
    L.x = act ? px[L.me] : 0.0;
    L.y = act ? py[L.me] : 0.0;
    L.z = act ? pz[L.me] : 0.0;
    L.p = act ? 0 : 0x7fffffff;
    L.cnt = 0;
    L.sx = L.sy = L.sz = 0.0;
    const unsigned long long gmask =
        (GS == 64) ? ~0ull : (((1ull << (GS & 63)) - 1ull) << (lane & ~(GS - 1)));
    int i = 0;
    while (true) {
        int ii = i;
        if constexpr (GS == 64) ii = __builtin_amdgcn_readfirstlane(i);
        const bool alive = ii < nnodes;
        if (!__any(alive)) break;
        bool desc = false, inter = false;
        int skip = ii + 1;
        if (alive) {
            const WNode nd = nodes[ii];
            skip = nd.skip;
            if (L.p == ii) {
                desc = walk_visit<PREC>(L, nd, ii, G, eps2, theta2);
                inter = L.inter;
            }
        }
        if (taping) tape_push(tape, b0 >> 6, lane, nrec, __ballot(inter), ii);
        const bool gdesc = (__ballot(desc) & gmask) != 0ull;
        const int nxt = gdesc ? ii + 1 : skip;
        i = nxt > ii ? nxt : ii + 1;  // strictly increasing: the loop always ends
        ++nit;
    }
    if (wo.cost && lane == 0) wo.cost[wv] = nit;
    if (act) {
        if constexpr (PREC == WALK_FP64_STRICT) {
            ax[L.me] = L.sx; ay[L.me] = L.sy; az[L.me] = L.sz;
        } else {
            ax[L.me] = G * L.sx; ay[L.me] = G * L.sy; az[L.me] = G * L.sz;
        }
        if (icount) icount[L.me] = L.cnt;
    }
    if (taping) tape_close(tape, b0 >> 6, lane, nrec);
    unsigned long long c = (unsigned long long)L.cnt;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) atomicAdd(itotal, c);
}

// Kahan step: add t to (s, k), k holding the negated running error.
__device__ __forceinline__ void kahan_add(float &s, float &k, float t) {
    const float y = t - k;
    const float u = s + y;
    k = (u - s) - y;
    s = u;
}

#define CSG_FOLD_EVERY 32

// FP32 walk over WNode32 records, same traversal as k_walk (groups of GS lanes).
// Offsets: d = fp32(C - x) with the subtraction in FP64 (one FP32 rounding of
// the offset).  Accumulation is compensated in blocks: each lane adds its terms
// in FP32 into a block sum, and every CSG_FOLD_EVERY node steps (a wave-uniform
// count) the block sums are Kahan-added into the running sums.  G multiplies
// the result.  k_walk32w is the default one-traversal-per-wavefront form; this
// grouped form serves GS = 8, 16, 32 (and 64 with an explicit DFS position).
template <int GS>
__global__ __launch_bounds__(CSG_TPB) void k_walk32(
    const WNode32 *__restrict__ nodes, const BuildParams *__restrict__ bp,
    const int *__restrict__ sidx, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    double G, double eps2, double theta2, int nb8, int chunk, TapeOut tape) {
    const int n_live = bp->n_live;
    const int nnodes = bp->n_nodes;
    if (bp->overflow) return;
    const int lane = threadIdx.x & 63;
    const int tile = xcd_tile(nb8, chunk);
    const int b0 = (tile * CSG_TPB + (int)(threadIdx.x & ~63));
    if (b0 >= n_live) return;
    const int b = b0 + lane;
    const bool act = b < n_live;
    const bool taping = (GS == 64) && tape.rec != nullptr;
    int nrec = 0;
    const int me = act ? sidx[b] : -1;
    const double x = act ? px[me] : 0.0, y = act ? py[me] : 0.0, z = act ? pz[me] : 0.0;
    const float eps2f = (float)eps2;  // theta is folded into the records' s2
    (void)theta2;
    const unsigned long long gmask =
        (GS == 64) ? ~0ull : (((1ull << (GS & 63)) - 1ull) << (lane & ~(GS - 1)));
    int p = act ? 0 : 0x7fffffff;
    int cnt = 0;
    float bx = 0.f, by = 0.f, bz = 0.f;               // block sums
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    int i = 0, steps = 0;
    while (true) {
        int ii = i;
        if constexpr (GS == 64) {
            ii = __builtin_amdgcn_readfirstlane(i);
            if (ii >= nnodes) break;
        } else {
            if (!__any(ii < nnodes)) break;
        }
        const bool alive = ii < nnodes;
        const WNode32 nd = nodes[alive ? ii : 0];
        // branch-free step: lanes whose DFS position is not ii compute and discard
        const bool here = alive && p == ii;
        const float dx = (float)(nd.cx - x);
        const float dy = (float)(nd.cy - y);
        const float dz = (float)(nd.cz - z);
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        // leaves carry s2 = -1, so crit holds for them and only self is skipped
        const bool crit = nd.s2 < d2;
        const bool inter = here && crit && nd.body != me;
        const bool desc = here && !crit;
        const float yv = __builtin_amdgcn_rsqf(d2 + eps2f);
        const float f = inter ? nd.m * (yv * yv * yv) : 0.0f;
        bx = fmaf(dx, f, bx);
        by = fmaf(dy, f, by);
        bz = fmaf(dz, f, bz);
        cnt += inter ? 1 : 0;
        if (taping) tape_push(tape, b0 >> 6, lane, nrec, __ballot(inter), ii);
        if (here) p = desc ? ii + 1 : nd.skip;
        if (++steps == CSG_FOLD_EVERY) {
            steps = 0;
            kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
            bx = by = bz = 0.f;
        }
        const bool gdesc = (__ballot(desc) & gmask) != 0ull;
        const int nxt = gdesc ? ii + 1 : nd.skip;
        i = (nxt > ii) ? nxt : ii + 1;  // strictly increasing: the loop always ends
    }
    kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
    if (act) {
        ax[me] = G * ((double)sx - (double)kx);
        ay[me] = G * ((double)sy - (double)ky);
        az[me] = G * ((double)sz - (double)kz);
        if (icount) icount[me] = cnt;
    }
    if (taping) tape_close(tape, b0 >> 6, lane, nrec);
    unsigned long long c = (unsigned long long)cnt;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) atomicAdd(itotal, c);
}

// Dispatch order of the walks, from each wave's node-step count in the previous
// walk (with its neighbours', k_cost_near).  The walk runs blocks of gw waves (gw consecutive waves: 64 gw sorted,
// spatially adjacent bodies that share a CU and its scalar cache); block b of
// the grid runs on XCD b % 8 (workgroups are dealt out to the XCDs
// round-robin), and wave slot s of the grid walks wave order[s].  The blocks'
// groups are cut into 8 contiguous ranges of sorted bodies, one per XCD, of
// about equal total step count: each XCD's L2 then serves one eighth of the
// space (the near-field nodes its waves share).  Each range is dispatched
// longest first (longest processing time first: long traversals start early
// and short ones fill the end, where an XCD otherwise drains with few waves per
// SIMD).  With split > 1 an XCD's range is cut further into split sub-ranges
// of equal step count, dispatched one after the other (each longest first), so
// that the waves running together on an XCD span a smaller part of the space.
// A boundary moves from the equal-count one by at most ng / 64 groups, so an
// XCD's range holds at most capg = WO_CAP(ng) groups; the slots past its end
// hold -1 (no wave).  Block r of 8 (1024 threads) orders XCD r's range by a
// counting sort on the group's step count quantized to 4096 levels, descending;
// the order inside a level is whatever the LDS atomics give, which changes only
// the timing.  Every block computes the same boundaries.
#define WO_XCDS 8
#define WO_REG 16
#define WO_MAX_SPLIT 4
#define WO_CAP(ng) ((((5 * (ng)) / 32 + 8) + 3) & ~3)
__device__ __forceinline__ int wo_cost(const int *__restrict__ cost, int w) {
    return min(max(cost[w], 0), 1 << 24);
}

// A wave's expected cost for the next dispatch order: the largest of its own and
// its neighbours' (WO_NEAR on each side) costs in the previous walk.  Between two
// walks the bodies move and the waves' membership shifts along the wave order, so
// an expensive (wide) group can land a wave or two away from where it was; taking
// the neighbours' costs keeps such a wave near the front of the dispatch (1M
// merger, list walk at theta 0.5, 100 substeps: 0.99 ms with two neighbours a
// side, 1.01 with one, 1.12 ms with each wave's own cost).
#define WO_NEAR 2
__global__ __launch_bounds__(CSG_TPB) void k_cost_near(const int *__restrict__ cost, int nw,
                                                       int *__restrict__ out) {
    const int w = blockIdx.x * CSG_TPB + threadIdx.x;
    if (w >= nw) return;
    int c = 0;
#pragma unroll
    for (int k = -WO_NEAR; k <= WO_NEAR; ++k)
        if (w + k >= 0 && w + k < nw) c = max(c, wo_cost(cost, w + k));
    out[w] = c;
}
__device__ __forceinline__ void wave_order_block(int r, const int *__restrict__ cost, int nw,
                                                 int gw, int split, int capg,
                                                 int *__restrict__ order) {
    __shared__ int hist[4096];
    __shared__ long long wsum[16];
    __shared__ int bnd[WO_XCDS * WO_MAX_SPLIT + 1];
    __shared__ int part[16];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int ng = (nw + gw - 1) / gw;
    const int R = WO_XCDS * split;   // ranges: XCD x takes ranges x split .. x split + split - 1
    // step count of group g (stale or garbage counts are clamped: any value
    // still yields a permutation)
    auto gcost = [&](int g) {
        int c = 0;
        for (int k = 0; k < gw; ++k) {
            const int w = g * gw + k;
            if (w < nw) c += wo_cost(cost, w);
        }
        return c;
    };
    // chunk sums of the group counts (thread t: groups [c0, c1)) and their prefix
    const int per = (ng + 1023) / 1024;
    const int c0 = min(ng, t * per), c1 = min(ng, c0 + per);
    // chunks of up to WO_REG one-wave groups (up to 1M bodies) are loaded at
    // once and kept in registers for the boundary search
    int cv[WO_REG];
    const bool regs = gw == 1 && per <= WO_REG;
    long long s = 0;
    if (regs) {
#pragma unroll
        for (int k = 0; k < WO_REG; ++k) cv[k] = c0 + k < c1 ? wo_cost(cost, c0 + k) : 0;
#pragma unroll
        for (int k = 0; k < WO_REG; ++k) s += cv[k];
    } else {
        for (int g = c0; g < c1; ++g) s += gcost(g);
    }
    long long inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[wid] = inc;
    if (t <= R) bnd[t] = (int)(((long long)t * ng) / R);   // equal counts
    __syncthreads();
    long long ex = inc - s, T = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < wid) ex += wsum[k];
        T += wsum[k];
    }
    // boundary k: one past the group whose prefix reaches k T / R (one chunk holds it)
    for (int k = 1; k < R; ++k) {
        const long long target = (T * k) / R;
        if (target > 0 && ex < target && ex + s >= target) {
            long long run = ex;
            int g = c0;
            if (regs) {
#pragma unroll
                for (int q = 0; q < WO_REG; ++q)
                    if (run < target) { run += cv[q]; g = c0 + q + 1; }
            } else {
                while (run < target) run += gcost(g++);
            }
            bnd[k] = g;
        }
    }
    __syncthreads();
    if (t == 0) {
        for (int k = 1; k < R; ++k) {
            const int mid = (int)(((long long)k * ng) / R), slack = ng / 64;
            const int bk = min(max(bnd[k], mid - slack), mid + slack);
            bnd[k] = min(max(bk, bnd[k - 1]), ng);
        }
    }
    __syncthreads();
    const int xlo = bnd[r * split], xhi = bnd[r * split + split];
    // position j of XCD r's list -> grid block 8 j + r, whose gw slots take the
    // group's waves in order
    auto place = [&](int j, int g) {
        if (j < capg)
            for (int k = 0; k < gw; ++k) {
                const int w = g * gw + k;
                order[(WO_XCDS * j + r) * gw + k] = (g >= 0 && w < nw) ? w : -1;
            }
    };
    for (int q = 0; q < split; ++q) {
        const int lo = bnd[r * split + q], hi = bnd[r * split + q + 1];
        for (int i = t; i < 4096; i += 1024) hist[i] = 0;
        // this thread's first group of the sub-range, loaded once
        const int myc = lo + t < hi ? gcost(lo + t) : 0;
        auto gc = [&](int g) { return g == lo + t ? myc : gcost(g); };
        int mx = 0;
        for (int g = lo + t; g < hi; g += 1024) mx = max(mx, gc(g));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
        if (lane == 0) part[wid] = mx;
        __syncthreads();
        mx = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) mx = max(mx, part[k]);
        int shift = 0;
        while ((mx >> shift) >= 4096) ++shift;
        auto bin = [&](int c) { return 4095 - (c >> shift); };
        for (int g = lo + t; g < hi; g += 1024) atomicAdd(&hist[bin(gc(g))], 1);
        __syncthreads();
        // exclusive scan of the 4096 bins: 4 per thread, then across threads
        int h[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) { h[k] = hist[4 * t + k]; sum += h[k]; }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        __syncthreads();
        if (lane == 63) part[wid] = incl;
        __syncthreads();
        int base0 = lo - xlo;   // the sub-range's first position in the XCD's list
        for (int k = 0; k < wid; ++k) base0 += part[k];
        int run = base0 + incl - sum;
#pragma unroll
        for (int k = 0; k < 4; ++k) { hist[4 * t + k] = run; run += h[k]; }
        __syncthreads();
        for (int g = lo + t; g < hi; g += 1024) place(atomicAdd(&hist[bin(gc(g))], 1), g);
        __syncthreads();
    }
    for (int j = (xhi - xlo) + t; j < capg; j += 1024) place(j, -1);
}

__global__ __launch_bounds__(1024) void k_wave_order(const int *__restrict__ cost, int nw, int gw,
                                                     int split, int capg, int *__restrict__ order) {
    wave_order_block(blockIdx.x, cost, nw, gw, split, capg, order);
}

// Between two walks of a substep run, in one launch of 1024-thread blocks: blocks
// 0 .. WO_XCDS-1 order the next walk's dispatch from the walk just finished
// (k_wave_order's work); the other blocks do k_kick_drift_bbox's update of the
// live bodies with one bounding-box partial each (part[8 (block - WO_XCDS) ..]),
// and k_zero_dead's zero accelerations of the dead bodies and zero interaction
// total for the next walk.
__global__ __launch_bounds__(1024, 8) void k_kick_drift_order(
    double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
    double *__restrict__ vx, double *__restrict__ vy, double *__restrict__ vz,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ m, int n, double dt, int kicks, double *__restrict__ part,
    int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    const int *__restrict__ cost, int nw, int gw, int split, int capg, int *__restrict__ order) {
    if (blockIdx.x < WO_XCDS) {
        wave_order_block(blockIdx.x, cost, nw, gw, split, capg, order);
        return;
    }
    const int bid = blockIdx.x - WO_XCDS, nkd = gridDim.x - WO_XCDS;
    if (bid == 0 && threadIdx.x == 0 && itotal) *itotal = 0ull;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double cnt = 0.0;
    for (int i = bid * 1024 + threadIdx.x; i < n; i += nkd * 1024) {
        if (!live_mass(m[i])) {
            ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0;
            if (icount) icount[i] = 0;
            continue;
        }
        double u = vx[i], v = vy[i], w = vz[i];
        const double hx = 0.5 * ax[i], hy = 0.5 * ay[i], hz = 0.5 * az[i];
        for (int k = 0; k < kicks; ++k) {
            u = u + hx * dt;
            v = v + hy * dt;
            w = w + hz * dt;
        }
        vx[i] = u; vy[i] = v; vz[i] = w;
        const double p[3] = {x[i] + u * dt, y[i] + v * dt, z[i] + w * dt};
        x[i] = p[0]; y[i] = p[1]; z[i] = p[2];
        cnt += 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], o, 64));
        }
        cnt += __shfl_xor(cnt, o, 64);
    }
    __shared__ double sw[16][7];
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) { sw[wid][a] = lo[a]; sw[wid][3 + a] = hi[a]; }
        sw[wid][6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int a = threadIdx.x;
        double v = sw[0][a];
        for (int k = 1; k < 16; ++k)
            v = a < 3 ? fmin(v, sw[k][a]) : (a < 6 ? fmax(v, sw[k][a]) : v + sw[k][a]);
        part[bid * 8 + a] = v;
    }
}

// Lane select on a wave mask held in SGPRs: m's bit for this lane ? t : f.
__device__ __forceinline__ float lane_sel_f32(unsigned long long m, float t, float f) {
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(f), "v"(t), "s"(m));
    return r;
}
__device__ __forceinline__ int lane_sel_i32(unsigned long long m, int t, int f) {
    int r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(f), "v"(t), "s"(m));
    return r;
}

// FP32 walk, one traversal per wavefront (k_walk32<64>'s traversal written on
// wave masks); same offsets, opening test, interaction and compensated
// accumulation as k_walk32.  The lane state is one integer p: a lane takes part
// in node ii while p <= ii, and accepting ii sets p = skip(ii).  A lane that
// descends keeps p, since the wave then moves to ii + 1 (any descending lane
// makes the whole wave descend); a waiting lane accepted a node J with ii inside
// J's subtree, so p = skip(J) >= skip(ii) and the wave cannot jump past it.
// Hence "p <= ii" holds exactly when the lane's own depth-first walk is at ii,
// as with the explicit position of k_walk32.
// Compares produce SGPR lane masks directly (the compare builtins), the
// participation / acceptance / descent logic runs on the scalar unit, and the
// two per-lane selects take those masks.
//   LANE_COUNTS: write per-body interaction counts (the wave total is always
//                kept, on the scalar unit);
//   TAPING:      record the walk tape;
//   OFF32:       the node array spans < 4 GiB, so records are addressed by a
//                32-bit byte offset (scalar loads with an SGPR offset);
//   TOTAL:       add the wave's interaction count to *itotal (the substep loop
//                counts only in its last walk, the one the statistics report);
//   SELF:        exclude the body's own leaf by index.  Without it (only for
//                softening > 0, no counts, no tape) the own leaf is taken as an
//                interaction: its offset is fp32(x - x) = 0 exactly and the
//                softened factor is finite, so it adds exactly 0.
// Wave w walks sorted bodies [64 w, 64 w + 64).  Without an order, w follows the
// XCD-chunked tile mapping; with one (WaveOrder), wave slot s of the grid takes
// w = order[s] (longest first, see k_wave_order).  Either way the results do not
// depend on the dispatch order.  With WaveOrder::cost set, every wave stores its
// node-step count there.
// The union traversal of one wavefront (k_walk32w's loop): lanes with act hold
// body me at (x, y, z); on return (sx - kx, sy - ky, sz - kz) is each lane's
// compensated sum without G, cnt its interaction count (LANE_COUNTS), wcnt the
// wave's (TOTAL), nrec its tape records (TAPING) and steps the wave's node-step
// count to within the last round.
template <bool LANE_COUNTS, bool TAPING, bool OFF32, bool TOTAL, bool SELF>
__device__ __forceinline__ void walk32_union(
    const WNode32 *__restrict__ nodes, int nnodes, int lane, int b0, bool act, int me, double x,
    double y, double z, float eps2f, const TapeOut &tape, int &nrec, int &cnt,
    unsigned long long &wcnt, float &sx, float &sy, float &sz, float &kx, float &ky, float &kz,
    int &steps) {
    const char *const base = (const char *)nodes;
    int p = act ? 0 : 0x7fffffff;
    unsigned wc32 = 0;
    float bx = 0.f, by = 0.f, bz = 0.f;               // block sums
    int ii = 0, quads = 0, folds = 0;
    // one node step of the wave (ii -> next node in the wave's depth-first order)
    auto step = [&]() {
        const WNode32 nd = OFF32
            ? *(const WNode32 *)(base + (unsigned)ii * (unsigned)sizeof(WNode32))
            : nodes[ii];
        const float dx = (float)(nd.cx - x);
        const float dy = (float)(nd.cy - y);
        const float dz = (float)(nd.cz - z);
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        // compare predicates: 41 = signed <=, 4 = ordered <, 33 = !=
        const unsigned long long H = __builtin_amdgcn_sicmp(p, ii, 41);
        // leaves carry s2 = -1, so the criterion holds for them
        const unsigned long long C = __builtin_amdgcn_fcmpf(nd.s2, d2, 4);
        const unsigned long long A = H & C;          // accepted (own leaf included)
        unsigned long long I = A;                    // interacting
        if constexpr (SELF) I &= __builtin_amdgcn_sicmp(nd.body, me, 33);
        const unsigned long long D = H & ~C;         // descending
        const float yv = __builtin_amdgcn_rsqf(d2 + eps2f);
        const float f = lane_sel_f32(I, nd.m * (yv * yv * yv), 0.0f);
        bx = fmaf(dx, f, bx);
        by = fmaf(dy, f, by);
        bz = fmaf(dz, f, bz);
        if constexpr (TOTAL) wc32 += (unsigned)__builtin_popcountll(I);   // <= 64 per step
        if constexpr (LANE_COUNTS) cnt += (int)((I >> lane) & 1ull);
        if constexpr (TAPING) tape_push(tape, b0 >> 6, lane, nrec, I, ii);
        p = lane_sel_i32(A, nd.skip, p);
        // strictly increasing even for a malformed skip: the loop always ends
        const int jump = max(nd.skip, ii + 1);
        ii = D ? ii + 1 : jump;
    };
    // eight steps per round and one end test per round: a round that starts
    // before the end may step onto up to seven of the inert padding records
    // after the last node (CSG_NODE_PAD); the compensated fold comes after every
    // CSG_FOLD_EVERY steps
    while (ii < nnodes) {
        step();
        step();
        step();
        step();
        step();
        step();
        step();
        step();
        if (++quads == CSG_FOLD_EVERY / 8) {
            quads = 0;
            ++folds;
            if constexpr (TOTAL) {
                wcnt += wc32;
                wc32 = 0;
            }
            kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
            bx = by = bz = 0.f;
        }
    }
    wcnt += wc32;
    steps = folds * CSG_FOLD_EVERY + quads * 8;
    kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
}

template <bool LANE_COUNTS, bool TAPING, bool OFF32, bool TOTAL, bool SELF>
__global__ __launch_bounds__(1024) void k_walk32w(
    const WNode32 *__restrict__ nodes, const BuildParams *__restrict__ bp,
    const int *__restrict__ sidx, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    double G, double eps2, double theta2, int nb8, int chunk, TapeOut tape, WaveOrder wo) {
    const int n_live = bp->n_live;
    const int nnodes = bp->n_nodes;
    if (bp->overflow || bp->order_bad) return;   // a failed build: no walk
    const int lane = threadIdx.x & 63;
    const int wv = walk_wave(nb8, chunk, wo);
    if (wv < 0) return;
    const int b0 = wv * 64;
    if (b0 >= n_live) {
        if (wo.cost && lane == 0 && wv < wo.n) wo.cost[wv] = 0;
        return;
    }
    const int b = b0 + lane;
    const bool act = b < n_live;
    int nrec = 0;
    const int me = act ? sidx[b] : -1;
    const double x = act ? px[me] : 0.0, y = act ? py[me] : 0.0, z = act ? pz[me] : 0.0;
    const float eps2f = (float)eps2;  // theta is folded into the records' s2
    (void)theta2;
    int cnt = 0, steps = 0;
    unsigned long long wcnt = 0;
    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    walk32_union<LANE_COUNTS, TAPING, OFF32, TOTAL, SELF>(nodes, nnodes, lane, b0, act, me, x, y,
        z, eps2f, tape, nrec, cnt, wcnt, sx, sy, sz, kx, ky, kz, steps);
    // node steps of this wave, to within the last round (the dispatch order's key)
    if (wo.cost && lane == 0) wo.cost[wv] = steps;
    if (act) {
        ax[me] = G * ((double)sx - (double)kx);
        ay[me] = G * ((double)sy - (double)ky);
        az[me] = G * ((double)sz - (double)kz);
        if constexpr (LANE_COUNTS) icount[me] = cnt;
    }
    if constexpr (TAPING) tape_close(tape, b0 >> 6, lane, nrec);
    if constexpr (TOTAL) {
        if (lane == 0) atomicAdd(itotal, wcnt);
    }
}

// FP32 walk on wave-shared interaction lists.  A wavefront holds 64 bodies
// consecutive in the walk order and treats them as one group: a node is
// accepted for the whole group when the opening criterion holds at the group's
// bounding box (s2 < dmin^2, dmin the distance from the node's centre of mass
// to the box), which implies the reference criterion (2h)^2/d^2 < theta^2 of
// src/octree.c:203-205 for every body of the group; leaves are always taken.
// The accepted nodes form the group's interaction list, and every lane sums
// the whole list for its body, so each list entry keeps all 64 lanes busy with
// one interaction.  Offsets are group-relative: the group's box centre g is
// formed in FP64, every entry holds fp32(C - g) and every lane fp32(x - g), both
// subtractions in FP64, and a lane's offset is their FP32 difference.
//
// Traversal: a stack in LDS of sibling chains (first node, end of the parent's
// subtree).  Each round pops up to 64 chains, one per lane; a lane tests its
// chain's first node, appends it to the list when accepted, and pushes the rest
// of the chain (the node's skip, while below the end) and, for an opened node,
// its children's chain (node + 1 up to its skip).  The buffered list (WL_LCAP
// entries) is summed by all lanes whenever the next round could overflow it,
// with a compensated fold every CSG_FOLD_EVERY entries.
//
// A group whose traversal would overrun the stack, or whose list grows past
// lmax entries (a wavefront whose bodies are spread over a wide box), falls
// back to the union traversal with each body's own criterion (walk32_union).
// With wmode (one word per wave, may be null), a wave that fell back takes the
// union traversal directly in its next `sticky` walks (its bodies change little
// from one substep to the next), then tries the list again: each walk is then a
// deterministic function of the state and of the walks since the bodies were
// set.  Both traversals apply at least the reference criterion to every body.
//   CNT:   2: per-body interaction counts into icount (when not null) and the
//          wave's total into *itotal, the own leaf not counted (by comparing
//          every entry's leaf body with the lane's own); 1: only the total, from
//          the list length (only for theta^2 < 1/3, where a body's own leaf is
//          always listed; spos maps the wave order to sorted positions (null:
//          the same) and L is the common-prefix array, for the bodies merged
//          into another's depth-64 leaf); 0: no counts;
//   SELF:  exclude the own leaf by index (softening below the FP32 range).
//          Without it the own leaf's offset is exactly 0 and its term is 0.
// WPB independent wavefronts per workgroup, each on its own LDS slice; wave slot
// s = WPB * blockIdx.x + (wavefront in block) takes wave order[s] with an order,
// else wave s.  With WaveOrder::cost set every wave stores a cost in
// node-step units (the dispatch order's key).  Waves that fell back add 1 to
// *fallbacks when it is not null.
#define WL_LCAP 128   // 256: 1% faster at theta 0.5 on the 1M merger, 6% slower at 0.7
#define WL_SCAP 384

__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Orders the LDS accesses of one wavefront's lanes (the wave's own slice of the
// workgroup's LDS; no other wavefront touches it).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CNT, bool SELF, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_walk32l(
    const WNode32 *__restrict__ nodes, const BuildParams *__restrict__ bp,
    const int *__restrict__ sidx, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    int *__restrict__ fallbacks, int *__restrict__ wmode, int sticky, const int *__restrict__ spos,
    const signed char *__restrict__ L, double G, double eps2, int lmax, WaveOrder wo) {
    constexpr bool FULL = CNT == 2, TOT = CNT >= 1;
    __shared__ int2 stk_all[WPB][WL_SCAP];
    __shared__ float4 lst_all[WPB][WL_LCAP];
    __shared__ int lbd_all[WPB][FULL || SELF ? WL_LCAP : 1];
    const int n_live = bp->n_live;
    const int nnodes = bp->n_nodes;
    if (bp->overflow || bp->order_bad) return;   // a failed build: no walk
    const int lane = threadIdx.x & 63;
    // the wavefront's index in its block, as a wave-uniform (scalar) value
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int2 *const stk = stk_all[wib];
    float4 *const lst = lst_all[wib];
    int *const lbd = lbd_all[wib];
    const int slot = (int)blockIdx.x * WPB + wib;
    int wv = slot;
    if (wo.order) {
        if (slot >= wo.slots) return;
        wv = wo.order[slot];
        if (wv < 0 || wv >= wo.n) return;
    }
    const int b0 = wv * 64;
    if (b0 >= n_live) {
        if (wo.cost && lane == 0 && wv < wo.n) wo.cost[wv] = 0;
        return;
    }
    const int b = b0 + lane;
    const bool act = b < n_live;
    const int me = act ? sidx[b] : -1;
    const double x = act ? px[me] : 0.0, y = act ? py[me] : 0.0, z = act ? pz[me] : 0.0;
    const float eps2f = (float)eps2;
    // the group's bounding box
    double lx = act ? x : DBL_MAX, ly = act ? y : DBL_MAX, lz = act ? z : DBL_MAX;
    double hx = act ? x : -DBL_MAX, hy = act ? y : -DBL_MAX, hz = act ? z : -DBL_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lx = fmin(lx, __shfl_xor(lx, o, 64)); hx = fmax(hx, __shfl_xor(hx, o, 64));
        ly = fmin(ly, __shfl_xor(ly, o, 64)); hy = fmax(hy, __shfl_xor(hy, o, 64));
        lz = fmin(lz, __shfl_xor(lz, o, 64)); hz = fmax(hz, __shfl_xor(hz, o, 64));
    }
    const double gx = 0.5 * (lx + hx), gy = 0.5 * (ly + hy), gz = 0.5 * (lz + hz);
    // half extents, widened by 2^-20 so the FP32 box holds every body
    const float wx = (float)(0.5 * (hx - lx)) * (1.0f + 0x1p-20f);
    const float wy = (float)(0.5 * (hy - ly)) * (1.0f + 0x1p-20f);
    const float wz = (float)(0.5 * (hz - lz)) * (1.0f + 0x1p-20f);
    const float qx = (float)(x - gx), qy = (float)(y - gy), qz = (float)(z - gz);

    float sx = 0.f, sy = 0.f, sz = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    float bx = 0.f, by = 0.f, bz = 0.f;   // block sums
    int blk = 0;                           // 8-entry blocks since the last fold
    int selfs = 0;                         // own-leaf entries seen (FULL)
    long long nent = 0;                    // list entries summed
    int rounds = 0;
    // sum the buffered list lst[0, nl) into every lane's block sums, in whole blocks of
    // 8 entries; the last call pads the rest with massless entries, the others move it
    // to the front of the list and return its length
    auto flush = [&](int nl, bool last) -> int {
        const int n8 = last ? (nl + 7) & ~7 : nl & ~7;
        if (last && lane < n8 - nl) {   // padding: massless entries far away add exactly 0
            lst[nl + lane] = make_float4(1e18f, 0.f, 0.f, 0.f);
            if constexpr (FULL || SELF) lbd[nl + lane] = -1;
        }
        wave_lds_sync();
        for (int q = 0; q < n8; q += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const float4 e = lst[q + u];
                const float dx = e.x - qx, dy = e.y - qy, dz = e.z - qz;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2f)));
                const float yv = __builtin_amdgcn_rsqf(r2);
                float f = e.w * (yv * yv * yv);
                if constexpr (FULL || SELF) {
                    const bool own = lbd[q + u] == me;
                    if constexpr (SELF) f = own ? 0.f : f;
                    if constexpr (FULL) selfs += own ? 1 : 0;
                }
                bx = fmaf(dx, f, bx);
                by = fmaf(dy, f, by);
                bz = fmaf(dz, f, bz);
            }
            if (++blk == CSG_FOLD_EVERY / 8) {
                blk = 0;
                kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
                bx = by = bz = 0.f;
            }
        }
        const int rest = last ? 0 : nl - n8;   // below 8 (n8 >= 64 here): no overlap
        nent += nl - rest;
        float4 e{};
        int bd = -1;
        if (lane < rest) {
            e = lst[n8 + lane];
            if constexpr (FULL || SELF) bd = lbd[n8 + lane];
        }
        wave_lds_sync();   // every lane has read the list before it is refilled
        if (lane < rest) {
            lst[lane] = e;
            if constexpr (FULL || SELF) lbd[lane] = bd;
        }
        return rest;
    };

    // walks left to take the union traversal
    const int mode = wmode ? __builtin_amdgcn_readfirstlane(wmode[wv]) : 0;
    if (lane == 0) stk[0] = make_int2(0, nnodes);
    int top = mode > 0 ? 0 : 1, nl = 0;
    bool fell = mode > 0;
    while (top > 0) {
        wave_lds_sync();
        ++rounds;
        const int take = top < 64 ? top : 64;
        const bool has = lane < take;
        const int2 ch = has ? stk[top - 1 - lane] : make_int2(0, 0);
        top -= take;
        const WNode32 nd = nodes[ch.x];
        const float rx = (float)(nd.cx - gx), ry = (float)(nd.cy - gy), rz = (float)(nd.cz - gz);
        const float tx = fmaxf(fabsf(rx) - wx, 0.f);
        const float ty = fmaxf(fabsf(ry) - wy, 0.f);
        const float tz = fmaxf(fabsf(rz) - wz, 0.f);
        const float dm2 = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
        // leaves carry s2 = -1: always taken
        const bool acc = has && nd.s2 < dm2;
        // chains start past the node that makes them, so the traversal always ends
        const bool opn = has && !acc && ch.x + 1 < nd.skip;
        const bool sib = has && nd.skip > ch.x && nd.skip < ch.y;
        const unsigned long long A = __ballot(acc), O = __ballot(opn), S = __ballot(sib);
        wave_lds_sync();   // every lane has read its chain before the stack is written
        if (acc) {
            const int k = nl + lane_rank(A);
            lst[k] = make_float4(rx, ry, rz, nd.m);
            if constexpr (FULL || SELF) lbd[k] = nd.body;
        }
        nl += __popcll(A);
        // the rest of each chain below the children, so children come off first
        if (sib) stk[top + lane_rank(S)] = make_int2(nd.skip, ch.y);
        top += __popcll(S);
        if (opn) stk[top + lane_rank(O)] = make_int2(ch.x + 1, nd.skip);
        top += __popcll(O);
        if (nl > WL_LCAP - 64 || top == 0) nl = flush(nl, top == 0);
        // a round pops at most 64 chains and pushes at most 128
        if (top > WL_SCAP - 64 || nent > lmax) {
            fell = true;
            break;
        }
    }
    int cnt = 0, steps = 0;
    unsigned long long wcnt = 0;
    if (fell) {
        sx = sy = sz = kx = ky = kz = 0.f;
        int nrec = 0;
        const TapeOut none{nullptr, nullptr, 0, nullptr};
        // counting excludes the own leaf by index, as the union walk's counting forms do
        walk32_union<FULL, false, false, TOT, SELF || TOT>(nodes, nnodes, lane, b0, act, me,
            x, y, z, eps2f, none, nrec, cnt, wcnt, sx, sy, sz, kx, ky, kz, steps);
        if (lane == 0) {
            if (mode == 0 && fallbacks) atomicAdd(fallbacks, 1);
            if (wmode) wmode[wv] = mode > 0 ? mode - 1 : sticky;
        }
    } else {
        kahan_add(sx, kx, bx); kahan_add(sy, ky, by); kahan_add(sz, kz, bz);
        if constexpr (FULL) {
            cnt = act ? (int)nent - selfs : 0;
            int t = cnt;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            wcnt = (unsigned long long)(unsigned)t;
        } else if constexpr (TOT) {
            // every lane sums every entry but its own leaf, and its own leaf is
            // listed (theta^2 < 1/3: no cell holding a body of the group passes the
            // group test), unless the body merged into another's depth-64 leaf
            const int pos = act ? (spos ? spos[b] : b) : 0;
            const bool merged = act && pos > 0 && L[pos - 1] >= 64;
            const int nact = __popcll(__ballot(act)), nmerged = __popcll(__ballot(merged));
            wcnt = (unsigned long long)nact * (unsigned long long)nent -
                   (unsigned long long)(nact - nmerged);
        }
    }
    // cost in node-step units: a list entry costs about half a union step
    if (wo.cost && lane == 0) wo.cost[wv] = (int)(fell ? steps + nent / 2 : nent / 2) + 3 * rounds;
    if (act) {
        ax[me] = G * ((double)sx - (double)kx);
        ay[me] = G * ((double)sy - (double)ky);
        az[me] = G * ((double)sz - (double)kz);
        if constexpr (FULL) {
            if (icount) icount[me] = cnt;
        }
    }
    if constexpr (TOT) {
        if (lane == 0) atomicAdd(itotal, wcnt);
    }
}

// 64-bit select on a wave mask held in SGPRs: m's bit for this lane ? t : f.
__device__ __forceinline__ double lane_sel_f64(unsigned long long m, double t, double f) {
    const long long tb = __double_as_longlong(t), fb = __double_as_longlong(f);
    const int lo = lane_sel_i32(m, (int)tb, (int)fb);
    const int hi = lane_sel_i32(m, (int)(tb >> 32), (int)(fb >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// FP64 walks, one traversal per wavefront on lane masks: k_walk32w's scheme
// (node index in a scalar register, records by scalar loads, participation,
// acceptance and descent as SGPR lane masks, four steps per round over the inert
// padding records after the tree) with walk_visit's per-lane arithmetic.  A
// lane visits exactly its own depth-first sequence in pre-order and adds the
// same terms in the same order as k_walk, so the accelerations and counts are
// k_walk's bit for bit.  Every lane evaluates the opening test and the force at
// every node of its wave's union; lanes outside the interacting mask add a
// zero term (x + 0 = x for the sums, which never hold -0).  The node array
// spans < 4 GiB (32-bit byte offsets).
template <int PREC, bool TAPING>
__global__ __launch_bounds__(1024) void k_walk64w(
    const WNode *__restrict__ nodes, const BuildParams *__restrict__ bp,
    const int *__restrict__ sidx, const double *__restrict__ px, const double *__restrict__ py,
    const double *__restrict__ pz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, int *__restrict__ icount, unsigned long long *__restrict__ itotal,
    double G, double eps2, double theta2, int nb8, int chunk, TapeOut tape, WaveOrder wo) {
    const int n_live = bp->n_live;
    const int nnodes = bp->n_nodes;
    if (bp->overflow) return;
    const int lane = threadIdx.x & 63;
    const int wv = walk_wave(nb8, chunk, wo);
    if (wv < 0) return;
    const int b0 = wv * 64;
    if (b0 >= n_live) {
        if (wo.cost && lane == 0 && wv < wo.n) wo.cost[wv] = 0;
        return;
    }
    const int b = b0 + lane;
    const bool act = b < n_live;
    int nrec = 0;
    const int me = act ? sidx[b] : -1;
    const double x = act ? px[me] : 0.0, y = act ? py[me] : 0.0, z = act ? pz[me] : 0.0;
    const char *const base = (const char *)nodes;
    int p = act ? 0 : 0x7fffffff;
    int cnt = 0;
    unsigned long long wcnt = 0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int ii = 0, quads = 0;
    auto step = [&]() {
        const WNode nd = *(const WNode *)(base + (unsigned)ii * (unsigned)sizeof(WNode));
        const double dx = nd.cx - x, dy = nd.cy - y, dz = nd.cz - z;
        double d2;
        if constexpr (PREC == WALK_FP64_STRICT) d2 = dx * dx + dy * dy + dz * dz;
        else d2 = fma(dx, dx, fma(dy, dy, dz * dz));
        // compare predicates: 41 = signed <=, 33 = !=
        const unsigned long long H = __builtin_amdgcn_sicmp(p, ii, 41);
        const bool leaf = nd.body >= 0;
        // internal node: the reference's opening test; a leaf interacts unless
        // it is the body itself
        const bool accept = open_accept_exact(nd.s2, d2, theta2);
        const unsigned long long C = leaf ? ~0ull : __ballot(accept);
        unsigned long long D = H & ~C;   // descending
        unsigned long long I = H & C;    // interacting
        if (leaf) I &= __builtin_amdgcn_sicmp(nd.body, me, 33);
        if constexpr (PREC == WALK_FP64_STRICT) {
            if (nd.m == 0.0) { D = 0ull; I = 0ull; }   // an empty cell is passed over
        }
        const unsigned long long V = H & ~D;   // lanes that move past the node
        const double r2 = d2 + eps2;
        if constexpr (PREC == WALK_FP64_STRICT) {
            // reference arithmetic and evaluation order, no contraction; the
            // correctly rounded divide and square root only on steps where a
            // lane interacts (a wave-uniform branch)
            if (I) {
                const double w = lane_sel_f64(I, 1.0 / (r2 * sqrt(r2)), 0.0);
                const double gm = G * nd.m;
                sx += gm * dx * w;
                sy += gm * dy * w;
                sz += gm * dz * w;
            }
        } else {
            const double yv = rsq64_walk(r2);
            const double f = lane_sel_f64(I, nd.m * (yv * yv * yv), 0.0);
            sx = fma(dx, f, sx);
            sy = fma(dy, f, sy);
            sz = fma(dz, f, sz);
        }
        wcnt += (unsigned long long)__builtin_popcountll(I);
        cnt += (int)((I >> lane) & 1ull);
        if constexpr (TAPING) tape_push(tape, b0 >> 6, lane, nrec, I, ii);
        p = lane_sel_i32(V, nd.skip, p);
        // strictly increasing even for a malformed skip: the loop always ends
        const int jump = max(nd.skip, ii + 1);
        ii = D ? ii + 1 : jump;
    };
    while (ii < nnodes) {
        step();
        step();
        step();
        step();
        ++quads;
    }
    if (wo.cost && lane == 0) wo.cost[wv] = quads * 4;
    if (act) {
        if constexpr (PREC == WALK_FP64_STRICT) {
            ax[me] = sx; ay[me] = sy; az[me] = sz;
        } else {
            ax[me] = G * sx; ay[me] = G * sy; az[me] = G * sz;
        }
        if (icount) icount[me] = cnt;
    }
    if constexpr (TAPING) tape_close(tape, b0 >> 6, lane, nrec);
    if (lane == 0) atomicAdd(itotal, wcnt);
}

// Dead bodies get a = 0 (src/octree.c:227-231).
__global__ __launch_bounds__(CSG_TPB) void k_zero_dead(const double *__restrict__ m, int n,
                                                       double *__restrict__ ax,
                                                       double *__restrict__ ay,
                                                       double *__restrict__ az,
                                                       int *__restrict__ icount,
                                                       unsigned long long *__restrict__ total) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i == 0 && total) *total = 0ull;   // the walk's interaction total starts here
    if (i >= n) return;
    if (!live_mass(m[i])) {
        ax[i] = 0.0; ay[i] = 0.0; az[i] = 0.0;
        if (icount) icount[i] = 0;
    }
}

// =================================================================== KDK
// v += (0.5*a)*dt; x += v*dt for live bodies (src/integrator.c:22-31).
__global__ __launch_bounds__(CSG_TPB) void k_kick_drift(double *__restrict__ x,
                                                        double *__restrict__ y,
                                                        double *__restrict__ z,
                                                        double *__restrict__ vx,
                                                        double *__restrict__ vy,
                                                        double *__restrict__ vz,
                                                        const double *__restrict__ ax,
                                                        const double *__restrict__ ay,
                                                        const double *__restrict__ az,
                                                        const double *__restrict__ m, int n,
                                                        double dt, int kicks) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n || !live_mass(m[i])) return;
    double u = vx[i], v = vy[i], w = vz[i];
    const double hx = 0.5 * ax[i], hy = 0.5 * ay[i], hz = 0.5 * az[i];
    for (int k = 0; k < kicks; ++k) {
        u = u + hx * dt;
        v = v + hy * dt;
        w = w + hz * dt;
    }
    vx[i] = u; vy[i] = v; vz[i] = w;
    x[i] = x[i] + u * dt;
    y[i] = y[i] + v * dt;
    z[i] = z[i] + w * dt;
}

// k_kick_drift over a grid-stride loop that also leaves k_bbox_partial's
// per-block bounds of the drifted positions in part (same layout), so that the
// build that follows starts from k_bbox_final.
__global__ __launch_bounds__(CSG_TPB) void k_kick_drift_bbox(
    double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
    double *__restrict__ vx, double *__restrict__ vy, double *__restrict__ vz,
    const double *__restrict__ ax, const double *__restrict__ ay, const double *__restrict__ az,
    const double *__restrict__ m, int n, double dt, int kicks, double *__restrict__ part) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double cnt = 0.0;
    for (int i = blockIdx.x * CSG_TPB + threadIdx.x; i < n; i += gridDim.x * CSG_TPB) {
        if (!live_mass(m[i])) continue;
        double u = vx[i], v = vy[i], w = vz[i];
        const double hx = 0.5 * ax[i], hy = 0.5 * ay[i], hz = 0.5 * az[i];
        for (int k = 0; k < kicks; ++k) {
            u = u + hx * dt;
            v = v + hy * dt;
            w = w + hz * dt;
        }
        vx[i] = u; vy[i] = v; vz[i] = w;
        const double p[3] = {x[i] + u * dt, y[i] + v * dt, z[i] + w * dt};
        x[i] = p[0]; y[i] = p[1]; z[i] = p[2];
        cnt += 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
    __shared__ double s[7][CSG_TPB];
    for (int a = 0; a < 3; ++a) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
    s[6][threadIdx.x] = cnt;
    __syncthreads();
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const int o = threadIdx.x + st;
            for (int a = 0; a < 3; ++a) {
                if (s[a][o] < s[a][threadIdx.x]) s[a][threadIdx.x] = s[a][o];
                if (s[3 + a][o] > s[3 + a][threadIdx.x]) s[3 + a][threadIdx.x] = s[3 + a][o];
            }
            s[6][threadIdx.x] += s[6][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x][0];
}

// v += (0.5*a)*dt for live bodies (src/integrator.c:45-51).
__global__ __launch_bounds__(CSG_TPB) void k_kick(double *__restrict__ vx,
                                                  double *__restrict__ vy,
                                                  double *__restrict__ vz,
                                                  const double *__restrict__ ax,
                                                  const double *__restrict__ ay,
                                                  const double *__restrict__ az,
                                                  const double *__restrict__ m, int n,
                                                  double dt) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n || !live_mass(m[i])) return;
    vx[i] = vx[i] + 0.5 * ax[i] * dt;
    vy[i] = vy[i] + 0.5 * ay[i] * dt;
    vz[i] = vz[i] + 0.5 * az[i] * dt;
}

// ============================================================= AoS <-> SoA
__global__ __launch_bounds__(CSG_TPB) void k_unpack(const Body *__restrict__ aos, int n,
                                                    double *x, double *y, double *z, double *vx,
                                                    double *vy, double *vz, double *ax,
                                                    double *ay, double *az, double *m) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    const Body &b = aos[i];
    x[i] = b.x; y[i] = b.y; z[i] = b.z;
    vx[i] = b.vx; vy[i] = b.vy; vz[i] = b.vz;
    ax[i] = b.ax; ay[i] = b.ay; az[i] = b.az;
    m[i] = b.mass;
}

__global__ __launch_bounds__(CSG_TPB) void k_unpack_acc(const Body *__restrict__ aos, int n,
                                                        double *ax, double *ay, double *az) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    ax[i] = aos[i].ax; ay[i] = aos[i].ay; az[i] = aos[i].az;
}

__global__ __launch_bounds__(CSG_TPB) void k_pack(Body *__restrict__ aos, int n,
                                                  const double *x, const double *y,
                                                  const double *z, const double *vx,
                                                  const double *vy, const double *vz,
                                                  const double *ax, const double *ay,
                                                  const double *az, const double *m) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    Body &b = aos[i];
    b.x = x[i]; b.y = y[i]; b.z = z[i];
    b.vx = vx[i]; b.vy = vy[i]; b.vz = vz[i];
    b.ax = ax[i]; b.ay = ay[i]; b.az = az[i];
    b.mass = m[i];
}

// ======================================================= reference-layout export
__global__ __launch_bounds__(CSG_TPB) void k_export_nodes(const WNode *__restrict__ nodes,
                                                          const uint64_t *__restrict__ key,
                                                          const uint64_t *__restrict__ ext,
                                                          const BuildParams *__restrict__ bp,
                                                          OctreeNode *__restrict__ pool) {
    const int v = blockIdx.x * CSG_TPB + threadIdx.x;
    if (v >= bp->n_nodes) return;
    const WNode w = nodes[v];
    OctreeNode o;
    o.cx = w.cx; o.cy = w.cy; o.cz = w.cz;
    o.total_mass = w.m;
    double cx = bp->root[0], cy = bp->root[1], cz = bp->root[2], h = bp->root[3];
    if (bp->n_live > 0)
        for (int l = 0; l < w.depth; ++l) cell_child(digit_at(key, ext, w.first, l), cx, cy, cz, h);
    o.center_x = cx; o.center_y = cy; o.center_z = cz;
    o.half_size = h;
    o.body_index = w.body;
    for (int c = 0; c < 8; ++c) o.children[c] = -1;
    pool[v] = o;
}

// Each non-root node writes its id into its parent's slot for its octant.
__global__ __launch_bounds__(CSG_TPB) void k_export_link(const WNode *__restrict__ nodes,
                                                         const uint64_t *__restrict__ key,
                                                         const uint64_t *__restrict__ ext,
                                                         const signed char *__restrict__ L,
                                                         const int *__restrict__ nbase,
                                                         const BuildParams *__restrict__ bp,
                                                         OctreeNode *__restrict__ pool) {
    const int v = blockIdx.x * CSG_TPB + threadIdx.x;
    if (v >= bp->n_nodes || v == 0 || bp->n_live == 0) return;
    const WNode w = nodes[v];
    const int b = w.first, d = w.depth;
    const int lpb = (b > 0) ? L[b - 1] : -1;
    int parent;
    if (v > nbase[b]) {
        parent = v - 1;  // nodes opened by one body are consecutive, depth + 1 each
    } else {
        const int bp0 = range_begin(key, ext, b, d - 1);
        const int lpp = (bp0 > 0) ? L[bp0 - 1] : -1;
        parent = nbase[bp0] + (d - 1) - (min(lpp, 63) + 1);
    }
    (void)lpb;
    if (parent < 0 || parent >= v) return;  // a parent precedes its children in pre-order
    pool[parent].children[digit_at(key, ext, b, d - 1)] = v;
}

// ================================================== walk of a foreign pool
// Exact reference traversal over a reference-layout pool (src/octree.c:171-218),
// FP64 with IEEE sqrt/divide, explicit stack, children in slot order.
#define CSG_POOL_STACK 600
__global__ __launch_bounds__(CSG_TPB) void k_walk_pool(const OctreeNode *__restrict__ pool,
                                                       int pool_size, int root,
                                                       const double *__restrict__ px,
                                                       const double *__restrict__ py,
                                                       const double *__restrict__ pz,
                                                       const double *__restrict__ pm, int n,
                                                       double *__restrict__ ax,
                                                       double *__restrict__ ay,
                                                       double *__restrict__ az,
                                                       int *__restrict__ icount, double G,
                                                       double eps2, double theta,
                                                       int *__restrict__ err) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int cnt = 0;
    if (live_mass(pm[i])) {
        const double x = px[i], y = py[i], z = pz[i];
        const double th2 = theta * theta;
        int stack[CSG_POOL_STACK];
        int top = 0;
        stack[top++] = root;
        while (top > 0) {
            const int v = stack[--top];
            if (v < 0 || v >= pool_size) { atomicOr(err, 2); continue; }
            const OctreeNode &nd = pool[v];
            if (nd.total_mass == 0.0) continue;
            const double dx = nd.cx - x, dy = nd.cy - y, dz = nd.cz - z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            bool inter;
            if (nd.body_index >= 0) inter = nd.body_index != i;
            else {
                const double s = 2.0 * nd.half_size;
                inter = d2 > 0.0 && (s * s) / d2 < th2;
                if (!inter) {
                    for (int c = 7; c >= 0; --c) {
                        const int ch = nd.children[c];
                        if (ch < 0) continue;
                        if (top >= CSG_POOL_STACK) { atomicOr(err, 1); break; }
                        stack[top++] = ch;
                    }
                    continue;
                }
            }
            if (inter) {
                const double r2 = d2 + eps2;
                const double w = 1.0 / (r2 * sqrt(r2));
                const double gm = G * nd.total_mass;
                sx += gm * dx * w;
                sy += gm * dy * w;
                sz += gm * dz * w;
                ++cnt;
            }
        }
    }
    ax[i] = sx; ay[i] = sy; az[i] = sz;
    if (icount) icount[i] = cnt;
}

// ================================================================ judges
// Direct-sum accelerations of targets idx[0..k) against source chunk
// blockIdx.y, one wave per 64 targets; chunk partials summed in chunk order by
// k_direct_reduce.  Arithmetic as tests/test_physics.c:54-71.
__global__ __launch_bounds__(64) void k_direct(const int *__restrict__ idx, int k,
                                               const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z,
                                               const double *__restrict__ m, int n, int chunk,
                                               double G, double eps2, double *__restrict__ part) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    const bool act = t < k;
    const int i = act ? idx[t] : 0;
    const double xi = x[i], yi = y[i], zi = z[i];
    const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double w = 1.0 / (r2 * sqrt(r2));
        const double gm = G * m[j];
        if (j != i) {
            sx += gm * dx * w;
            sy += gm * dy * w;
            sz += gm * dz * w;
        }
    }
    if (act) {
        double *o = part + ((size_t)blockIdx.y * k + t) * 3;
        o[0] = sx; o[1] = sy; o[2] = sz;
    }
}

__global__ __launch_bounds__(CSG_TPB) void k_direct_reduce(const double *__restrict__ part, int k,
                                                           int nchunk, double *__restrict__ out) {
    const int t = blockIdx.x * CSG_TPB + threadIdx.x;
    if (t >= k) return;
    for (int a = 0; a < 3; ++a) {
        double s = 0.0;
        for (int c = 0; c < nchunk; ++c) s += part[((size_t)c * k + t) * 3 + a];
        out[t * 3 + a] = s;
    }
}

// Potential energy partials: sum over j != i of G m_i m_j / sqrt(r^2 + eps^2)
// for the 256 targets of a block against source chunk blockIdx.y.
__global__ __launch_bounds__(CSG_TPB) void k_potential(const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       const double *__restrict__ m, int n,
                                                       int chunk, double G, double eps2,
                                                       double *__restrict__ part) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    const bool act = i < n;
    const int ii = act ? i : 0;
    const double xi = x[ii], yi = y[ii], zi = z[ii], mi = act ? m[ii] : 0.0;
    const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    double s = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        const double r = sqrt(dx * dx + dy * dy + dz * dz + eps2);
        if (j != ii) s += G * mi * m[j] / r;
    }
    __shared__ double red[CSG_TPB];
    red[threadIdx.x] = act ? s : 0.0;
    __syncthreads();
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Kinetic energy and momentum per block: out[blk*4 + {0: KE, 1..3: p}].
__global__ __launch_bounds__(CSG_TPB) void k_kinetic(const double *__restrict__ vx,
                                                     const double *__restrict__ vy,
                                                     const double *__restrict__ vz,
                                                     const double *__restrict__ m, int n,
                                                     double *__restrict__ part) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    if (i < n) {
        const double mm = m[i], u = vx[i], w = vy[i], q = vz[i];
        v[0] = 0.5 * mm * (u * u + w * w + q * q);
        v[1] = mm * u; v[2] = mm * w; v[3] = mm * q;
    }
    __shared__ double red[4][CSG_TPB];
    for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = CSG_TPB / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// Fixed-order compensated sum of cnt values with stride `stride` starting at
// offset `off` (one block).
__global__ __launch_bounds__(CSG_TPB) void k_sum(const double *__restrict__ v, int cnt,
                                                 int stride, int off, double *__restrict__ out) {
    __shared__ double hs[CSG_TPB], ls[CSG_TPB];
    double s = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < cnt; k += CSG_TPB) {
        const double t = v[(size_t)k * stride + off] - c;
        const double u = s + t;
        c = (u - s) - t;
        s = u;
    }
    hs[threadIdx.x] = s;
    ls[threadIdx.x] = -c;
    __syncthreads();
    if (threadIdx.x == 0) {
        DD acc{0.0, 0.0};
        for (int k = 0; k < CSG_TPB; ++k) acc = dd_add(acc, dd_add(DD{hs[k], 0.0}, DD{ls[k], 0.0}));
        *out = acc.h;
    }
}
