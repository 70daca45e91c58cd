// csg_core.hip - context management and the csg_* C-ABI (include/csg.h) of the
// HIP gravity core.  Every entry point makes the context's device current on the
// calling thread, launches on the context's stream and returns 0 or a negative
// COSMOSIM_E_* code.  Kernels live in csg_device.h.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "csg_device.h"
#include "cosmosim_ext.h"
#include "csg.h"

#define CSG_EV_MAX 4096
enum { PH_BUILD = 0, PH_WALK = 1, PH_KDK = 2, PH_NB = 3, PH_SPH = 4 };

struct csg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cap = 0;       // body capacity
    int n = 0;         // bodies held
    int node_cap = 0;  // node records (8 per body, as the reference pool)
    // state, caller's body order
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr,
           *vz = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *m = nullptr;
    Body *aos = nullptr;  // last uploaded records (untouched fields travel back from here)
    // build
    BuildParams *bp = nullptr;
    BuildParams *hbp = nullptr;  // pinned mirror
    double *bbox_part = nullptr;
    uint64_t *key = nullptr, *key2 = nullptr, *ext = nullptr, *ext_tmp = nullptr;
    int *idx = nullptr, *idx2 = nullptr;
    int *widx = nullptr;   // FP32 wave-mask walk: the sorted bodies in Peano-Hilbert order
    int *hpos = nullptr;   // ... and their sorted positions (hpos[r]: widx[r] = idx[hpos[r]])
    int *rs_hist = nullptr, *rs_tot = nullptr;
    RsCtl *rs_ctl = nullptr;           // body sort: tickets and digit counts
    uint64_t *rs_status = nullptr;     // body sort: per-tile digit counts (look-back)
    unsigned rs_epoch = 0;             // epoch of the last sort pass
    double *xs = nullptr, *ys = nullptr, *zs = nullptr, *ms = nullptr;  // sorted copies
    signed char *lcp = nullptr;
    signed char *lmin1 = nullptr, *lmin2 = nullptr;  // lcp minima per 64 / 4096 sorted bodies
    int *cnt = nullptr, *nbase = nullptr, *scan_part = nullptr;
    DD *pref = nullptr, *dd_part = nullptr;
    DD *tpref = nullptr;   // build: each thread's exclusive prefix inside its tile
    WNode *nodes = nullptr;
    WNode32 *nodes32 = nullptr;
    int *nlast = nullptr;   // build: each node's last sorted member (Peano-Hilbert pass)
    int *icount = nullptr;
    unsigned long long *itotal = nullptr;
    int tree_valid = 0;
    int fail_seen = 0;         // BuildParams::fail_count as of the last overflow check,
                               // including the build that check examined
    int debug_fail_build = 0;  // test hook: the n-th build from now runs with no node
                               // capacity (overflows like a tree that does not fit)
    // export / foreign pool / diagnostics
    OctreeNode *pool_dev = nullptr;
    size_t pool_dev_cap = 0;
    int *err_dev = nullptr;
    void *diag = nullptr;
    size_t diag_cap = 0;
    // timing: event pairs recorded on the stream without host waits and read
    // back in collect_timers(); totals accumulate until timing is re-enabled
    int timing = 0;
    hipEvent_t *ev = nullptr;  // 2 * CSG_EV_MAX events, created on first use
    int ev_phase[CSG_EV_MAX];
    int ev_count = 0, ev_open = -1;
    double ms_build = 0, ms_walk = 0, ms_kdk = 0, ms_nb = 0, ms_sph = 0;
    long long nb_queries = 0;
    int timed_walks = 0;
    // adjoint: walk tape, transposed records, per-body adjoint buffers, checkpoints
    int tape_on = 0, tape_cap = 0, tape_waves = 0, tape_valid = 0;
    int tape_wmap = 0;   // the tape's wavefronts hold the bodies of hpos, not sorted runs
    double tape_G = 1.0, tape_eps2 = 0.0;
    TapeRec *tape = nullptr;
    int *tcount = nullptr, *toff = nullptr, *tape_total = nullptr, *tape_ovf = nullptr;
    size_t rec_cap = 0;
    uint32_t *rkey = nullptr, *rkey2 = nullptr;   // adjoint records' node ids
    int *rval = nullptr, *rval2 = nullptr, *rhist = nullptr, *rtot = nullptr;
    double *rsum = nullptr, *ws = nullptr, *xbar_s = nullptr, *Dv = nullptr, *Uv = nullptr;
    double *cbar = nullptr, *xb = nullptr, *vb = nullptr, *ab = nullptr, *gtmp = nullptr;
    DD *vpart = nullptr;
    WNode *nodes_alt = nullptr;
    int *seg_first = nullptr, *seg_end = nullptr, *seg_long = nullptr, *seg_nlong = nullptr;
    unsigned long long *u64tmp = nullptr;
    double *ckpt = nullptr, *seg = nullptr;
    int ckpt_slots = 0, seg_slots = 0;
    int rec_valid = 0, rec_k = 0, rec_every = 1, rec_prec = 0;
    struct FrozenStep *frozen = nullptr;
    int n_frozen = 0;
    double rec_dt = 0, rec_G = 1, rec_eps = 0, rec_theta = 0;
    double theta_hint = 0.5;  // opening angle of the next walks (set by every walk entry)
    double th32 = -1.0;       // opening angle folded into the FP32 records (-1: none)
    // FP64 walks: 0 = wave-mask traversal, else lanes per traversal group of the
    // grouped walk (8, 16, 32 or 64); -1 (default) = wave-mask for FP64 and
    // groups of 8 for FP64-strict, whose correctly rounded divide and square
    // root make each step dear (1M-body merger walk: FP64 3.36 ms wave-mask vs
    // 3.71 grouped; strict 5.42 vs 4.45)
    int walk_group = -1;
    int walk_group32 = 0;   // FP32 walk: 0 = wave-mask traversal, else lanes per group
    // Untaped walks take their bodies in Peano-Hilbert order (hilbert_waves) when it
    // pays: the grouped FP64 walks from hilbert_min bodies (below, the three launches
    // of the order cost more than the shorter traversals save), the wave-mask walks
    // only with more wavefronts than the GPU holds at once (resident_waves): when
    // every wavefront is resident the longest one sets the walk's time, and Hilbert
    // order can make the longest one longer (a group of bodies spread through a
    // sparse region); beyond that the total of the traversals counts, which it cuts.
    int hilbert_waves = 1;
    int hilbert_min = 16384;
    int resident_waves = 8192;   // compute units x 4 SIMDs x 8 wavefronts
    int widx_valid = 0;          // widx holds the current tree's Peano-Hilbert order
    // the kick-drift before the next walk already wrote its dispatch order (for these
    // waves per block, slots per XCD and waves), zeroed the dead bodies'
    // accelerations and the interaction total
    int order_ready = 0, order_gw = 0, order_capg = 0, order_nw = 0, order_split = 0;
    int counts_valid = 0;   // icount holds the per-body counts of the last walk
    int nodes64_valid = 0;  // the FP64 node records belong to the current tree
    double4 *xyzm = nullptr;  // build scratch: (x, y, z, m) interleaved, body order
    double4 *geom = nullptr;  // neighbour queries: cell centre and half size per node
    size_t geom_cap = 0;
    int geom_valid = 0;       // geom describes the current tree
    int *gq = nullptr;        // SPH: live gas bodies in tree order, tile offsets, count
    size_t gq_cap = 0;
    int gq_valid = 0;         // gq describes the current tree
    void *snap = nullptr;     // snapshot records and scan scratch
    size_t snap_cap = 0;
    void *qbuf = nullptr;     // quasar: flags, ranks, black-hole list, draws
    size_t qbuf_cap = 0;
    void *qtmp = nullptr;     // quasar: compaction scratch
    size_t qtmp_cap = 0;
    struct QDev *qdev = nullptr, *qhost = nullptr;   // quasar step state (device, pinned)
    unsigned long long qrng = 1;   // quasar random stream (xorshift64)
    int qphase = 0;                // quasar knot phase
    void *nbuf = nullptr;     // neighbour queries: scratch
    size_t nbuf_cap = 0;
    int *wcost = nullptr;   // FP32 wave-mask walk: node steps of each wave in the last walk
    int *worder = nullptr;  // dispatch order built from wcost (longest first)
    int wcost_n = 0;        // number of waves wcost describes (0: none)
    int wave_order = 1;     // dispatch the wave-mask walk longest first
    int xcd_chunk = 16;     // walk tiles per XCD chunk (0: one contiguous range per XCD)
    int walk_block_waves = 1;  // waves per block of the ordered walks
    int wave_split = 2;        // sub-ranges per XCD range of the walk's dispatch order
                               // (1M merger: 1.74 / 1.68 / 1.72 / 1.71 ms walk for 1-4)
    int list_max = 4096;       // list entries after which a list walk's wave falls back, at
                               // theta 0.5 (scaled by 0.25 / theta^2 in launch_walk32l)
    int list_block_waves = 2;  // list walk: independent wavefronts per workgroup (1, 2, 4;
                               // 1M merger at theta 0.5: 1.35 / 1.29 / 1.31 ms walk)
    int *wl_fall = nullptr;    // list walks: waves that fell back (device counter)
    int *wmode = nullptr;      // list walks: per wave, walks left on the union traversal
    int *wcost_s = nullptr;    // the costs a dispatch order reads (k_cost_near of wcost)
    int list_wave_split = 1;   // sub-ranges per XCD range of the list walk's dispatch order
                               // (1M merger, theta 0.5: 1.27 / 1.33 / 1.40 ms for 1 / 2 / 3)
    int list_sticky = 256;     // walks a wave that fell back stays on the union traversal
                               // (0: every walk tries the list; results then depend only
                               // on the state, not on the walks before).  1M merger over
                               // 100 substeps, limit 4096: 0.947 / 0.942 / 0.939 ms walk
                               // for 64 / 256 / 100000
};

#define CSG_CHECK(expr)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "cosmosim: %s failed: %s\n", #expr, hipGetErrorString(e_));      \
            return COSMOSIM_E_HIP;                                                           \
        }                                                                                    \
    } while (0)

#define CSG_LAUNCHED()                                                                       \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) {                                                              \
            fprintf(stderr, "cosmosim: kernel launch failed: %s\n", hipGetErrorString(e_));  \
            return COSMOSIM_E_HIP;                                                           \
        }                                                                                    \
    } while (0)

static void adj_free(csg_ctx *c);  // csg_adjoint.inl
static void frozen_free(csg_ctx *c);
static void nb_free(csg_ctx *c);   // csg_neighbors.inl
static void sph_free(csg_ctx *c);  // csg_sph.inl
static void snap_free(csg_ctx *c); // csg_snapshot.inl
static void q_free(csg_ctx *c);    // csg_quasar.inl

static inline int blocks_for(long long n, int per) { return (int)((n + per - 1) / per); }
static inline int at_least1(int v) { return v < 1 ? 1 : v; }

template <class T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <class T>
static bool dalloc(T *&p, size_t count) {
    p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void **)&p, count * sizeof(T)) == hipSuccess;
}

static void free_buffers(csg_ctx *c) {
    dfree(c->x); dfree(c->y); dfree(c->z); dfree(c->vx); dfree(c->vy); dfree(c->vz);
    dfree(c->ax); dfree(c->ay); dfree(c->az); dfree(c->m); dfree(c->aos);
    dfree(c->bbox_part); dfree(c->key); dfree(c->key2); dfree(c->ext); dfree(c->ext_tmp);
    dfree(c->idx); dfree(c->idx2); dfree(c->widx); dfree(c->hpos); dfree(c->nlast); dfree(c->rs_hist); dfree(c->rs_tot); dfree(c->rs_ctl); dfree(c->rs_status); dfree(c->xs); dfree(c->ys); dfree(c->zs); dfree(c->ms);
    dfree(c->lcp); dfree(c->lmin1); dfree(c->lmin2); dfree(c->cnt); dfree(c->nbase); dfree(c->scan_part);
    dfree(c->pref); dfree(c->dd_part); dfree(c->tpref); dfree(c->nodes); dfree(c->nodes32); dfree(c->icount);
    dfree(c->wcost); dfree(c->worder); dfree(c->xyzm); dfree(c->wmode); dfree(c->wcost_s);
    c->wcost_n = 0;
    c->cap = 0;
    c->node_cap = 0;
}

template <class T>
__global__ __launch_bounds__(CSG_TPB) void k_clear(T *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * CSG_TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * CSG_TPB)
        p[i] = T(0);
}

template <class T>
static void clear_words(T *p, size_t n, hipStream_t st) {
    if (n == 0) return;
    const size_t b = (n + CSG_TPB - 1) / CSG_TPB;
    k_clear<T><<<(unsigned)(b < 4096 ? b : 4096), CSG_TPB, 0, st>>>(p, n);
}

// Chunk minima of the common-prefix depths: k_lcp writes 4 per block of
// CSG_TPB bodies (one per wavefront of 64), k_iscan_down one per 64 of those;
// padded so 16-byte reads past the last entry stay inside.
static constexpr size_t lmin1_count(size_t cap) { return 4 * ((cap + 1 + CSG_TPB - 1) / CSG_TPB); }
static constexpr size_t lmin1_bytes(size_t cap) { return (lmin1_count(cap) + 63) / 64 * 64 + 16; }
static constexpr size_t lmin2_bytes(size_t cap) { return (lmin1_count(cap) + 63) / 64 + 32; }

// Device buffers for `cap` bodies; about 900 bytes per body (the node records,
// 8 per body at 64 B, dominate; 4096 extra records cover deep chains of
// near-coincident bodies in small inputs).
static int alloc_buffers(csg_ctx *c, int cap) {
    const size_t C = (size_t)(cap < 1 ? 1 : cap);
    const size_t C1 = C + 1;
    const int nb = at_least1(blocks_for((long long)C1, CSG_TILE));
    bool ok = dalloc(c->x, C) && dalloc(c->y, C) && dalloc(c->z, C) && dalloc(c->vx, C) &&
              dalloc(c->vy, C) && dalloc(c->vz, C) && dalloc(c->ax, C) && dalloc(c->ay, C) &&
              dalloc(c->az, C) && dalloc(c->m, C) && dalloc(c->aos, C) &&
              dalloc(c->bbox_part, (size_t)CSG_BBOX_BLOCKS * 8) && dalloc(c->key, C) &&
              dalloc(c->key2, C) && dalloc(c->ext, 3 * C) && dalloc(c->ext_tmp, 3 * C) &&
              dalloc(c->idx, C) && dalloc(c->idx2, C) && dalloc(c->widx, C) &&
              dalloc(c->hpos, C) && dalloc(c->rs_hist, (size_t)256 * nb) &&
              dalloc(c->rs_tot, 256) && dalloc(c->xs, C) && dalloc(c->ys, C) && dalloc(c->zs, C) &&
              dalloc(c->ms, C) &&
              dalloc(c->lcp, C1 + 64) && dalloc(c->lmin1, lmin1_bytes(C)) &&
              dalloc(c->lmin2, lmin2_bytes(C)) && dalloc(c->cnt, C1) && dalloc(c->nbase, C1) &&
              dalloc(c->scan_part, (size_t)nb) && dalloc(c->pref, 4 * C1) &&
              dalloc(c->dd_part, (size_t)4 * nb) && dalloc(c->tpref, (size_t)4 * CSG_TPB * nb) && dalloc(c->nodes, 8 * C + 4096) && dalloc(c->nodes32, 8 * C + 4096) && dalloc(c->nlast, 8 * C + 4096) &&
              dalloc(c->icount, C) && dalloc(c->wcost, C / 64 + 1) && dalloc(c->wcost_s, C / 64 + 1) && dalloc(c->wmode, C / 64 + 1) && dalloc(c->worder, 2 * (size_t)(C / 64 + 1) + 4096) &&
              dalloc(c->xyzm, C) && dalloc(c->rs_ctl, 1) &&
              dalloc(c->rs_status, (size_t)256 * blocks_for((long long)C, RS_TILE));
    if (!ok) {
        (void)hipGetLastError();
        free_buffers(c);
        return COSMOSIM_E_ALLOC;
    }
    // look-back words start unpublished; the index buffers start as valid indices
    clear_words(c->rs_status, (size_t)256 * blocks_for((long long)C, RS_TILE), c->stream);
    clear_words(c->idx, C, c->stream);
    clear_words(c->idx2, C, c->stream);
    clear_words(c->widx, C, c->stream);
    clear_words(c->wmode, C / 64 + 1, c->stream);
    c->rs_epoch = 0;
    if (hipGetLastError() != hipSuccess) {
        free_buffers(c);
        return COSMOSIM_E_HIP;
    }
    c->cap = (int)C;
    c->node_cap = (int)(8 * C + 4096);
    return COSMOSIM_OK;
}

static int use(csg_ctx *c) {
    if (!c) return COSMOSIM_E_ARG;
    CSG_CHECK(hipSetDevice(c->device));
    return COSMOSIM_OK;
}

#define CSG_USE(c)                          \
    do {                                    \
        int s_ = use(c);                    \
        if (s_ != COSMOSIM_OK) return s_;   \
    } while (0)

static int ensure_diag(csg_ctx *c, size_t bytes) {
    if (bytes <= c->diag_cap) return COSMOSIM_OK;
    if (c->diag) (void)hipFree(c->diag);
    c->diag = nullptr;
    c->diag_cap = 0;
    if (hipMalloc(&c->diag, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return COSMOSIM_E_ALLOC;
    }
    c->diag_cap = bytes;
    return COSMOSIM_OK;
}

static void timer_start(csg_ctx *c) {
    if (!c->timing || !c->ev || c->ev_count >= CSG_EV_MAX) return;
    c->ev_open = c->ev_count;
    (void)hipEventRecord(c->ev[2 * c->ev_open], c->stream);
}
static void timer_stop(csg_ctx *c, int phase) {
    if (c->ev_open < 0) return;
    (void)hipEventRecord(c->ev[2 * c->ev_open + 1], c->stream);
    c->ev_phase[c->ev_open] = phase;
    c->ev_count = c->ev_open + 1;
    c->ev_open = -1;
}
// Wait for the recorded events and add their durations to the phase totals.
static void collect_timers(csg_ctx *c) {
    if (!c->ev || c->ev_count == 0) return;
    (void)hipEventSynchronize(c->ev[2 * c->ev_count - 1]);
    for (int k = 0; k < c->ev_count; ++k) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]);
        if (c->ev_phase[k] == PH_BUILD) c->ms_build += ms;
        else if (c->ev_phase[k] == PH_WALK) { c->ms_walk += ms; c->timed_walks += 1; }
        else if (c->ev_phase[k] == PH_NB) c->ms_nb += ms;
        else if (c->ev_phase[k] == PH_SPH) c->ms_sph += ms;
        else c->ms_kdk += ms;
    }
    c->ev_count = 0;
}

static int read_params(csg_ctx *c) {
    CSG_CHECK(hipMemcpyAsync(c->hbp, c->bp, sizeof(BuildParams), hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

extern "C" {

int csg_device_count(int *count) {
    if (!count) return COSMOSIM_E_ARG;
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return COSMOSIM_E_NO_DEVICE;
    }
    *count = k;
    return COSMOSIM_OK;
}

int csg_device_info(int device, char *name, int len, int *compute_units) {
    int k = 0;
    if (csg_device_count(&k) != COSMOSIM_OK || device < 0 || device >= k)
        return COSMOSIM_E_NO_DEVICE;
    hipDeviceProp_t p;
    CSG_CHECK(hipGetDeviceProperties(&p, device));
    if (name && len > 0) {
        snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    return COSMOSIM_OK;
}

int csg_device_memory(int device, unsigned long long *free_bytes, unsigned long long *total_bytes) {
    int k = 0;
    if (csg_device_count(&k) != COSMOSIM_OK || device < 0 || device >= k)
        return COSMOSIM_E_NO_DEVICE;
    CSG_CHECK(hipSetDevice(device));
    size_t f = 0, t = 0;
    CSG_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return COSMOSIM_OK;
}

int csg_create(int device, int capacity, csg_ctx **out) {
    if (!out || capacity < 0) return COSMOSIM_E_ARG;
    *out = nullptr;
    int k = 0;
    if (csg_device_count(&k) != COSMOSIM_OK || k == 0 || device < 0 || device >= k)
        return COSMOSIM_E_NO_DEVICE;
    csg_ctx *c = new (std::nothrow) csg_ctx();
    if (!c) return COSMOSIM_E_ALLOC;
    c->device = device;
    if (const char *g = getenv("COSMOSIM_WALK_GROUP")) {
        const int v = atoi(g);
        if (v == 0 || v == 8 || v == 16 || v == 32 || v == 64) c->walk_group = v;
    }
    if (const char *g = getenv("COSMOSIM_XCD_CHUNK")) c->xcd_chunk = atoi(g);
    if (const char *g = getenv("COSMOSIM_WAVE_ORDER")) c->wave_order = atoi(g) != 0;
    if (const char *g = getenv("COSMOSIM_HILBERT_WAVES")) c->hilbert_waves = atoi(g) != 0;
    if (const char *g = getenv("COSMOSIM_HILBERT_MIN_BODIES")) c->hilbert_min = atoi(g);
    if (const char *g = getenv("COSMOSIM_WAVE_SPLIT")) {
        const int v = atoi(g);
        if (v >= 1 && v <= WO_MAX_SPLIT) c->wave_split = v;
    }
    if (const char *g = getenv("COSMOSIM_WALK_BLOCK_WAVES")) {
        const int v = atoi(g);
        if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) c->walk_block_waves = v;
    }
    if (const char *g = getenv("COSMOSIM_LIST_MAX")) c->list_max = atoi(g);
    if (const char *g = getenv("COSMOSIM_LIST_STICKY")) c->list_sticky = atoi(g) > 0 ? atoi(g) : 0;
    if (const char *g = getenv("COSMOSIM_LIST_WAVE_SPLIT")) {
        const int v = atoi(g);
        if (v >= 1 && v <= WO_MAX_SPLIT) c->list_wave_split = v;
    }
    if (const char *g = getenv("COSMOSIM_LIST_BLOCK_WAVES")) {
        const int v = atoi(g);
        if (v == 1 || v == 2 || v == 4) c->list_block_waves = v;
    }
    if (const char *g = getenv("COSMOSIM_WALK_GROUP32")) {
        const int v = atoi(g);
        if (v == 0 || v == 8 || v == 16 || v == 32 || v == 64) c->walk_group32 = v;
    }
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return COSMOSIM_E_HIP;
    }
    if (hipMalloc((void **)&c->bp, sizeof(BuildParams)) != hipSuccess ||
        hipHostMalloc((void **)&c->hbp, sizeof(BuildParams), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&c->itotal, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void **)&c->wl_fall, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&c->err_dev, sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        csg_destroy(c);
        return COSMOSIM_E_ALLOC;
    }
    memset(c->hbp, 0, sizeof(BuildParams));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        c->resident_waves = 32 * cus;
    (void)hipGetLastError();
    if (const char *g = getenv("COSMOSIM_HILBERT_MIN_WAVES")) c->resident_waves = atoi(g);
    // the build counters start at zero (k_bbox_final counts earlier failures)
    clear_words((int *)c->bp, sizeof(BuildParams) / sizeof(int), c->stream);
    clear_words(c->wl_fall, 1, c->stream);
    int s = alloc_buffers(c, capacity);
    if (s != COSMOSIM_OK) {
        csg_destroy(c);
        return s;
    }
    *out = c;
    return COSMOSIM_OK;
}

void csg_destroy(csg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_buffers(c);
    adj_free(c);
    frozen_free(c);
    nb_free(c);
    sph_free(c);
    snap_free(c);
    q_free(c);
    dfree(c->bp);
    dfree(c->itotal);
    dfree(c->err_dev);
    dfree(c->wl_fall);
    dfree(c->pool_dev);
    if (c->diag) (void)hipFree(c->diag);
    if (c->hbp) (void)hipHostFree(c->hbp);
    if (c->ev) {
        for (int k = 0; k < 2 * CSG_EV_MAX; ++k)
            if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
        free(c->ev);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipGetLastError();
    delete c;
}

// Grow to `capacity` bodies, keeping the body state.
int csg_reserve(csg_ctx *c, int capacity) {
    CSG_USE(c);
    if (capacity < 0) return COSMOSIM_E_ARG;
    if (capacity <= c->cap) return COSMOSIM_OK;
    const int keep = c->n;
    double *old[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->m};
    Body *old_aos = c->aos;
    // detach the state arrays so free_buffers leaves them alone
    c->x = c->y = c->z = c->vx = c->vy = c->vz = c->ax = c->ay = c->az = c->m = nullptr;
    c->aos = nullptr;
    CSG_CHECK(hipStreamSynchronize(c->stream));
    free_buffers(c);
    int s = alloc_buffers(c, capacity);
    if (s == COSMOSIM_OK && keep > 0) {
        double *nw[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->m};
        for (int a = 0; a < 10; ++a)
            if (hipMemcpyAsync(nw[a], old[a], sizeof(double) * keep, hipMemcpyDeviceToDevice,
                               c->stream) != hipSuccess)
                s = COSMOSIM_E_HIP;
        if (hipMemcpyAsync(c->aos, old_aos, sizeof(Body) * keep, hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess)
            s = COSMOSIM_E_HIP;
        (void)hipStreamSynchronize(c->stream);
    }
    for (int a = 0; a < 10; ++a) dfree(old[a]);
    dfree(old_aos);
    c->tree_valid = 0;
    c->counts_valid = 0;
    if (s != COSMOSIM_OK) c->n = 0;
    return s;
}

int csg_size(const csg_ctx *c) { return c ? c->n : 0; }

int csg_set_bodies(csg_ctx *c, const Body *bodies, int n) {
    CSG_USE(c);
    if (n < 0 || (n > 0 && !bodies)) return COSMOSIM_E_ARG;
    int s = csg_reserve(c, n);
    if (s != COSMOSIM_OK) return s;
    c->n = n;
    c->tree_valid = 0;
    c->counts_valid = 0;
    c->wcost_n = 0;
    clear_words(c->wmode, (size_t)(c->cap / 64 + 1), c->stream);   // new bodies: no history
    if (n == 0) return COSMOSIM_OK;
    CSG_CHECK(hipMemcpyAsync(c->aos, bodies, sizeof(Body) * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
    k_unpack<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->aos, n, c->x, c->y, c->z, c->vx,
                                                               c->vy, c->vz, c->ax, c->ay, c->az,
                                                               c->m);
    CSG_LAUNCHED();
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_set_accelerations(csg_ctx *c, const Body *bodies, int n) {
    CSG_USE(c);
    if (n != c->n || (n > 0 && !bodies)) return COSMOSIM_E_ARG;
    if (n == 0) return COSMOSIM_OK;
    CSG_CHECK(hipMemcpyAsync(c->aos, bodies, sizeof(Body) * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
    k_unpack_acc<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->aos, n, c->ax, c->ay, c->az);
    CSG_LAUNCHED();
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_get_bodies(csg_ctx *c, Body *bodies, int n) {
    CSG_USE(c);
    if (n < 0 || n > c->n || (n > 0 && !bodies)) return COSMOSIM_E_ARG;
    if (n == 0) return COSMOSIM_OK;
    k_pack<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->aos, n, c->x, c->y, c->z, c->vx,
                                                             c->vy, c->vz, c->ax, c->ay, c->az, c->m);
    CSG_LAUNCHED();
    CSG_CHECK(hipMemcpyAsync(bodies, c->aos, sizeof(Body) * (size_t)n, hipMemcpyDeviceToHost,
                             c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

__global__ void k_interleave3(const double *a, const double *b, const double *cc, int n,
                              double *out) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = a[i];
    out[3 * (size_t)i + 1] = b[i];
    out[3 * (size_t)i + 2] = cc[i];
}
__global__ void k_deinterleave3(const double *in, int n, double *a, double *b, double *cc) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    a[i] = in[3 * (size_t)i];
    b[i] = in[3 * (size_t)i + 1];
    cc[i] = in[3 * (size_t)i + 2];
}
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_set_i32(int *p, int v) { *p = v; }

static int get3(csg_ctx *c, const double *a, const double *b, const double *cc, double *host) {
    const int n = c->n;
    k_interleave3<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, c->stream>>>(a, b, cc, n, (double *)c->diag);
    CSG_LAUNCHED();
    CSG_CHECK(hipMemcpyAsync(host, c->diag, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                             c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

static int set3(csg_ctx *c, double *a, double *b, double *cc, const double *host) {
    const int n = c->n;
    CSG_CHECK(hipMemcpyAsync(c->diag, host, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
    k_deinterleave3<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, c->stream>>>((const double *)c->diag, n,
                                                                      a, b, cc);
    CSG_LAUNCHED();
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_get_state(csg_ctx *c, double *pos, double *vel, double *acc, double *mass) {
    CSG_USE(c);
    if (c->n == 0) return COSMOSIM_OK;
    int s = ensure_diag(c, sizeof(double) * 3 * (size_t)c->n);
    if (s) return s;
    if (pos && (s = get3(c, c->x, c->y, c->z, pos))) return s;
    if (vel && (s = get3(c, c->vx, c->vy, c->vz, vel))) return s;
    if (acc && (s = get3(c, c->ax, c->ay, c->az, acc))) return s;
    if (mass) {
        CSG_CHECK(hipMemcpyAsync(mass, c->m, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
        CSG_CHECK(hipStreamSynchronize(c->stream));
    }
    return COSMOSIM_OK;
}

int csg_set_state(csg_ctx *c, const double *pos, const double *vel, const double *acc,
                  const double *mass) {
    CSG_USE(c);
    if (c->n == 0) return COSMOSIM_OK;
    int s = ensure_diag(c, sizeof(double) * 3 * (size_t)c->n);
    if (s) return s;
    if (pos && (s = set3(c, c->x, c->y, c->z, pos))) return s;
    if (vel && (s = set3(c, c->vx, c->vy, c->vz, vel))) return s;
    if (acc && (s = set3(c, c->ax, c->ay, c->az, acc))) return s;
    if (mass) {
        CSG_CHECK(hipMemcpyAsync(c->m, mass, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
        CSG_CHECK(hipStreamSynchronize(c->stream));
    }
    if (pos || mass) {
        c->tree_valid = 0;
        clear_words(c->wmode, (size_t)(c->cap / 64 + 1), c->stream);   // no list-walk history
    }
    return COSMOSIM_OK;
}

// Walk arithmetic codes the core accepts, and those whose walks read the FP32
// node records (the FP64 records are then emitted only on demand).
static inline bool prec_valid(int p) {
    return p >= COSMOSIM_PREC_FP64_STRICT && p <= COSMOSIM_PREC_FP32_LIST;
}
static inline bool prec_rec32(int p) {
    return p == COSMOSIM_PREC_FP32 || p == COSMOSIM_PREC_FP32_LIST;
}

static int launch_build(csg_ctx *c, bool with64 = true, int bbox = 0, int order = -1);

// Whether the untaped walk at `precision` over the current bodies takes the
// Peano-Hilbert wavefront order (so the build before it computes the order): the
// grouped FP64 walks from hilbert_min bodies, the wave-mask walks only with more
// wavefronts than the GPU holds at once; the grouped FP32 walk keeps sorted order.
static int hilbert_order_for(const csg_ctx *c, int precision) {
    // the list walk's groups are most compact in Peano-Hilbert order at every size
    if (precision == COSMOSIM_PREC_FP32_LIST) return c->hilbert_waves;
    if (!c->hilbert_waves || c->n < c->hilbert_min) return 0;
    if (precision == COSMOSIM_PREC_FP32 && c->walk_group32 != 0) return 0;
    const int g64 = c->walk_group >= 0 ? c->walk_group
                                       : (precision == COSMOSIM_PREC_FP64_STRICT ? 8 : 0);
    if (precision != COSMOSIM_PREC_FP32 && g64 != 0) return 1;
    return blocks_for(c->n, 64) > c->resident_waves ? 1 : 0;
}

// The FP64 node records of the current tree, re-emitted by a full build when the
// last build wrote only the FP32 records (same positions, so the same tree).
static int ensure_nodes64(csg_ctx *c, int precision) {
    if (prec_rec32(precision) || c->nodes64_valid) return COSMOSIM_OK;
    return launch_build(c, true, 0, hilbert_order_for(c, precision));
}

// Blocks of the bounding-box partials (k_bbox_partial's grid).
static int bbox_blocks(int n) {
    return n > 0 ? (blocks_for(n, CSG_TPB) < CSG_BBOX_BLOCKS ? blocks_for(n, CSG_TPB) : CSG_BBOX_BLOCKS)
                 : 1;
}

// Whether the next walk at `precision` (taped or not) over nw waves dispatches
// in longest-first order from the previous walk's step counts, and then its waves
// per block and slots per XCD range.
static bool walk_order_params(const csg_ctx *c, int precision, bool taped, int nw, int *gw,
                              int *capg, int *split) {
    const bool wave_mask32 = precision == COSMOSIM_PREC_FP32 && (c->walk_group32 == 0 || taped);
    const bool grouped32 = precision == COSMOSIM_PREC_FP32 && !wave_mask32;
    if (grouped32 || !c->wave_order || c->wcost_n != nw) return false;
    // blocks of gw waves (one by default: the finest balance; larger blocks, whose
    // waves share a CU's scalar cache, measured slower).  The FP64 walks' kernels are
    // bounded at CSG_TPB threads; the list walk's blocks hold list_block_waves
    // independent wavefronts.
    if (precision == COSMOSIM_PREC_FP32_LIST) *gw = c->list_block_waves;
    else *gw = wave_mask32 ? c->walk_block_waves : min(c->walk_block_waves, CSG_TPB / 64);
    *split = precision == COSMOSIM_PREC_FP32_LIST ? c->list_wave_split : c->wave_split;
    *capg = WO_CAP(blocks_for(nw, *gw));
    return true;
}

// The costs the next dispatch order reads: each wave's last cost with its
// neighbours' (k_cost_near), or the plain costs where there is no buffer for them.
static const int *order_costs(csg_ctx *c, int nw) {
    if (!c->wcost_s) return c->wcost;
    k_cost_near<<<blocks_for(nw, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->wcost, nw, c->wcost_s);
    return c->wcost_s;
}

// bbox (may be null): also leave the drifted positions' bounding-box partials
// in c->bbox_part and their block count in *bbox, for the build that follows.
// next_prec >= 0 (with bbox): the precision of the untaped walk that follows the
// build; when that walk dispatches in longest-first order, the order is computed
// in the same launch (k_kick_drift_order) and the walk skips its own.
static int launch_kick_drift(csg_ctx *c, double dt, int kicks, int *bbox = nullptr,
                             int next_prec = -1) {
    if (bbox) *bbox = 0;
    if (c->n == 0) return COSMOSIM_OK;
    const int nw = blocks_for(c->n, 64);
    int gw = 1, capg = 0, split = 1;
    if (bbox && next_prec >= 0 && walk_order_params(c, next_prec, false, nw, &gw, &capg, &split)) {
        const int nkd = min(blocks_for(c->n, 1024), CSG_BBOX_BLOCKS);
        *bbox = nkd;
        const int *oc = order_costs(c, nw);
        k_kick_drift_order<<<WO_XCDS + nkd, 1024, 0, c->stream>>>(c->x, c->y, c->z, c->vx,
            c->vy, c->vz, c->ax, c->ay, c->az, c->m, c->n, dt, kicks, c->bbox_part, c->icount,
            c->itotal, oc, nw, gw, split, capg, c->worder);
        c->order_ready = 1;
        c->order_gw = gw;
        c->order_split = split;
        c->order_capg = capg;
        c->order_nw = nw;
    } else if (bbox) {
        *bbox = bbox_blocks(c->n);
        k_kick_drift_bbox<<<*bbox, CSG_TPB, 0, c->stream>>>(c->x, c->y, c->z, c->vx, c->vy, c->vz,
            c->ax, c->ay, c->az, c->m, c->n, dt, kicks, c->bbox_part);
    } else {
        k_kick_drift<<<blocks_for(c->n, CSG_TPB), CSG_TPB, 0, c->stream>>>(
            c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->m, c->n, dt, kicks);
    }
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

static int launch_kick(csg_ctx *c, double dt) {
    if (c->n == 0) return COSMOSIM_OK;
    k_kick<<<blocks_for(c->n, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->vx, c->vy, c->vz, c->ax, c->ay,
                                                                c->az, c->m, c->n, dt);
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

// Whole build on the stream; no host synchronisation.
// with64: also write the FP64 node records (the FP64 walks, the adjoint and the
// tree export read them; the FP32 walk reads only the 40-byte records).
// bbox: the block count of bounding-box partials of the current positions that
// the preceding kick-drift left in c->bbox_part (0: compute them here).
// order: 1 / 0 compute the Peano-Hilbert wavefront order or not (hilbert_order_for
// of the walk that follows); -1 (a build with no particular walk in view): from
// hilbert_min bodies.
static int launch_build(csg_ctx *c, bool with64, int bbox, int order) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const int nbb = bbox > 0 ? bbox : bbox_blocks(n);
    int node_cap = c->node_cap;
    if (c->debug_fail_build > 0 && --c->debug_fail_build == 0) node_cap = 0;
    if (bbox <= 0) k_bbox_partial<<<nbb, CSG_TPB, 0, st>>>(c->x, c->y, c->z, c->m, n, c->bbox_part);
    k_bbox_final<<<1, CSG_TPB, 0, st>>>(c->bbox_part, nbb, c->bp, node_cap, c->rs_ctl);
    CSG_LAUNCHED();
    if (n > 0) {
        const int nt = blocks_for(n, RS_TILE);
        // the keys start in the buffer that makes the last pass end in key / idx
        constexpr bool odd = CSG_SORT_PASSES % 2 != 0;
        uint64_t *kin = odd ? c->key2 : c->key, *kout = odd ? c->key : c->key2;
        int *vin = odd ? c->idx2 : c->idx, *vout = odd ? c->idx : c->idx2;
        k_keys<<<nt, RS_TPB, 0, st>>>(c->x, c->y, c->z, c->m, n, c->bp, kin, c->xyzm,
                                      c->rs_ctl, c->scan_part, blocks_for(n + 1, CSG_TILE));
        if (c->rs_epoch > 0x7fffff00u) {   // epochs are 31 bits: start the words over
            clear_words(c->rs_status, (size_t)256 * blocks_for(c->cap, RS_TILE), st);
            c->rs_epoch = 0;
        }
        for (int pass = 0; pass < CSG_SORT_PASSES; ++pass) {
            const unsigned ep = ++c->rs_epoch;
            if (pass == 0)
                k_rs_sweep<true><<<nt, RS_TPB, 0, st>>>(kin, nullptr, kout, vout, n, pass, c->rs_ctl,
                                                         c->rs_status, ep, c->bp);
            else
                k_rs_sweep<false><<<nt, RS_TPB, 0, st>>>(kin, vin, kout, vout, n, pass, c->rs_ctl,
                                                          c->rs_status, ep, c->bp);
            uint64_t *tk = kin; kin = kout; kout = tk;
            int *tv = vin; vin = vout; vout = tv;
        }
        CSG_LAUNCHED();
        // the sorted pairs end in key / idx
        k_runs<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, st>>>(c->key, c->idx, c->x, c->y, c->z, c->bp, c->ext,
                                        c->idx2, c->ext_tmp);
        CSG_LAUNCHED();
    }
    const int n1 = n + 1;
    const int nlb = blocks_for(n1, CSG_TPB);
    // k_lcp also writes the sorted bodies' Peano-Hilbert keys into key2 (the sort's
    // spare key buffer)
    k_lcp<<<nlb, CSG_TPB, 0, st>>>(c->key, c->ext, c->bp, c->lcp, c->cnt, c->scan_part,
                                   c->lmin1, c->key2);
    const int nt = blocks_for(n1, CSG_TILE);
    k_iscan_down<<<nt, CSG_TPB, 0, st>>>(c->cnt, c->scan_part, c->bp, c->nbase, c->lmin1,
                                         4 * nlb, c->lmin2);
    k_gather_dd<<<nt, CSG_TPB, 0, st>>>(c->idx, c->xyzm, c->bp, c->xs, c->ys, c->zs, c->ms,
                                        c->dd_part, c->tpref);
    // k_dd_down2 also clears cnt for k_emit's Peano-Hilbert offsets
    k_dd_down2<<<nt, CSG_TPB, 0, st>>>(c->xs, c->ys, c->zs, c->ms, c->bp, c->dd_part, c->tpref,
                                       c->pref, c->cnt);
    const double inv_th2 = c->theta_hint > 0.0 ? 1.0 / (c->theta_hint * c->theta_hint) : HUGE_VAL;
    k_emit<<<at_least1(blocks_for(n, CSG_TPB)), CSG_TPB, 0, st>>>(
        c->key, c->ext, c->idx, c->lcp, c->lmin1, c->lmin2, c->nbase, c->pref, c->cap + 1, c->xs,
        c->ys, c->zs, c->ms, c->bp, with64 ? c->nodes : nullptr, c->nodes32, inv_th2, c->nlast);
    // the Peano-Hilbert wavefront order: the child-boundary steps go to the tile
    // prefixes' buffer, which k_dd_down2 is done with
    c->widx_valid = order >= 0 ? order : (c->hilbert_waves && n >= c->hilbert_min);
    if (c->widx_valid) {
        int *const hbnd = (int *)c->tpref;
        k_hilbert<<<at_least1(blocks_for(n, CSG_TPB)), CSG_TPB, 0, st>>>(
            c->lcp, c->nbase, c->nlast, c->key2, c->bp, c->cnt, hbnd);
        k_hsum<<<nt, CSG_TPB, 0, st>>>(c->cnt, hbnd, c->lcp, c->bp, c->scan_part);
        k_hrank<<<nt, CSG_TPB, 0, st>>>(c->cnt, hbnd, c->lcp, c->scan_part, c->idx, c->bp,
                                        c->widx, c->hpos);
    }
    CSG_LAUNCHED();
    c->tree_valid = 1;
    c->geom_valid = 0;
    c->gq_valid = 0;
    c->nodes64_valid = with64 ? 1 : 0;
    c->th32 = c->theta_hint;
    c->tape_valid = 0;   // a tape describes the tree it was recorded on
    return COSMOSIM_OK;
}

extern "C++" {
// FP64 walks: dispatched through `wo` (grid = blocks for the ordered slots) or
// the XCD-chunked mapping (grid = nb8); the grouped FP32 forms ignore the order.
// sidx64: the FP64 walks' body order (each lane sums its own terms in FP64, so the
// grouping into wavefronts changes nothing in the results; Peano-Hilbert order
// shortens the traversals); the grouped FP32 walk keeps sorted order.
template <int GS>
static void launch_walk_gs(csg_ctx *c, int precision, const int *sidx64, double G, double eps2,
                           double th2, int grid, int nb8, TapeOut tape, WaveOrder wo, int tpb) {
    hipStream_t st = c->stream;
    if (precision == COSMOSIM_PREC_FP64_STRICT)
        k_walk<WALK_FP64_STRICT, GS><<<grid, tpb, 0, st>>>(c->nodes, c->bp, sidx64, c->x, c->y,
            c->z, c->ax, c->ay, c->az, c->icount, c->itotal, G, eps2, th2, nb8, c->xcd_chunk, tape,
            wo);
    else if (precision == COSMOSIM_PREC_FP64)
        k_walk<WALK_FP64, GS><<<grid, tpb, 0, st>>>(c->nodes, c->bp, sidx64, c->x, c->y, c->z,
            c->ax, c->ay, c->az, c->icount, c->itotal, G, eps2, th2, nb8, c->xcd_chunk, tape, wo);
    else
        k_walk32<GS><<<nb8, CSG_TPB, 0, st>>>(c->nodes32, c->bp, c->idx, c->x, c->y, c->z, c->ax, c->ay,
                                          c->az, c->icount, c->itotal, G, eps2, th2, nb8, c->xcd_chunk, tape);
}

template <bool LC, bool TP, bool O32, bool TOT, bool SELF>
static void launch_walk32w(csg_ctx *c, const int *sidx, double G, double eps2, double th2,
                           int grid, int nb8, TapeOut tape, WaveOrder wo, int tpb) {
    k_walk32w<LC, TP, O32, TOT, SELF><<<grid, tpb, 0, c->stream>>>(c->nodes32, c->bp, sidx, c->x,
        c->y, c->z, c->ax, c->ay, c->az, c->icount, c->itotal, G, eps2, th2, nb8, c->xcd_chunk,
        tape, wo);
}

// cnt: 0 no counts, 1 the total only (from the list lengths), 2 per-body counts
// and the total; selfx: the own leaf excluded by index.  The fallback limit is
// list_max at theta 0.5 and scales as theta^-2: the lists grow as the opening
// angle shrinks, and so does the union traversal a long list competes with (1M
// merger over 100 substeps, walk ms for limits 3072 / 4096 / 6144: theta 0.4
// 2.17 / 1.62 / 1.38, theta 0.5 1.15 / 0.94 / 0.99; theta 0.7 with 2090 / 3135:
// 0.516 / 0.518).
template <int WPB>
static void launch_walk32l(csg_ctx *c, int cnt, bool selfx, const int *sidx, const int *spos,
                           double G, double eps2, double th2, int grid, WaveOrder wo) {
    int *ic = cnt == 2 ? c->icount : nullptr;
    const double lm = (double)c->list_max * (0.25 / fmax(th2, 1e-12));
    const int lmax = lm >= 2147483647.0 ? 2147483647 : (int)lm;
#define CSG_WL(C, S)                                                                         \
    k_walk32l<C, S, WPB><<<grid, 64 * WPB, 0, c->stream>>>(c->nodes32, c->bp, sidx, c->x, c->y, \
        c->z, c->ax, c->ay, c->az, ic, c->itotal, c->wl_fall, c->list_sticky ? c->wmode : nullptr, \
        c->list_sticky, spos, c->lcp, G, eps2, lmax, wo)
    if (selfx) {
        if (cnt == 2) CSG_WL(2, true); else if (cnt == 1) CSG_WL(1, true); else CSG_WL(0, true);
    } else {
        if (cnt == 2) CSG_WL(2, false); else if (cnt == 1) CSG_WL(1, false); else CSG_WL(0, false);
    }
#undef CSG_WL
}

template <int PREC, bool TP>
static void launch_walk64w(csg_ctx *c, const int *sidx, double G, double eps2, double th2,
                           int grid, int nb8, TapeOut tape, WaveOrder wo, int tpb) {
    k_walk64w<PREC, TP><<<grid, tpb, 0, c->stream>>>(c->nodes, c->bp, sidx, c->x, c->y, c->z,
        c->ax, c->ay, c->az, c->icount, c->itotal, G, eps2, th2, nb8, c->xcd_chunk, tape, wo);
}
}  // extern "C++"

// lane_counts: record per-body interaction counts (c->icount); walks inside the
// substep loops skip them.  count_total: keep the interaction total (the FP32
// substep loop keeps it only in its last walk; the statistics report that one).
// hilbert: the FP32 wave-mask walk takes its wavefronts' bodies in Peano-Hilbert
// order (c->widx) when the tree has one and there are more wavefronts than the GPU
// holds at once; false keeps sorted order.  A taped walk follows the same rule,
// so its forces are bit-identical to the untaped walk's (the recording, the
// checkpoint recompute and the time reversal follow one trajectory), and its tape
// then maps wavefront lanes to sorted bodies through c->hpos (c->tape_wmap).
static int launch_walk(csg_ctx *c, int precision, double G, double softening, double theta,
                       TapeOut tape = TapeOut{nullptr, nullptr, 0, nullptr},
                       bool lane_counts = true, bool count_total = true, bool hilbert = true) {
    const int n = c->n;
    hipStream_t st = c->stream;
    const double eps2 = softening * softening, th2 = theta * theta;
    c->theta_hint = theta;
    if (prec_rec32(precision) && c->tree_valid &&
        (c->th32 != theta ||
         (precision == COSMOSIM_PREC_FP32_LIST && hilbert && hilbert_order_for(c, precision) &&
          !c->widx_valid))) {
        // the FP32 records carry the opening angle: emit them again for this one;
        // the list walk's groups also need the Peano-Hilbert order, which a build
        // with no walk in view computes only from hilbert_min bodies
        const int s = launch_build(c, c->nodes64_valid != 0, 0, hilbert_order_for(c, precision));
        if (s) return s;
    }
    const bool pre = c->order_ready != 0;   // the kick-drift before did the order and zeroing
    c->order_ready = 0;
    if (n == 0) {
        k_set_u64<<<1, 1, 0, st>>>(c->itotal, 0ull);
        return COSMOSIM_OK;
    }
    if (!pre)
        k_zero_dead<<<blocks_for(n, CSG_TPB), CSG_TPB, 0, st>>>(c->m, n, c->ax, c->ay, c->az,
                                                                c->icount, c->itotal);
    int nb8 = blocks_for(n, CSG_TPB);
    const int gran = c->xcd_chunk > 0 ? 8 * c->xcd_chunk : 8;
    nb8 = (nb8 + gran - 1) / gran * gran;
    if (!prec_valid(precision) || (tape.rec && precision == COSMOSIM_PREC_FP32_LIST))
        return COSMOSIM_E_ARG;
    c->counts_valid = 0;
    const bool wave_mask32 = precision == COSMOSIM_PREC_FP32 && (c->walk_group32 == 0 || tape.rec);
    const bool grouped32 = precision == COSMOSIM_PREC_FP32 && !wave_mask32;
    // longest-first dispatch from the previous walk's per-wave step counts (the
    // wave-mask FP32 walk and the FP64 walks record them)
    const int nw = blocks_for(n, 64);
    WaveOrder wo{nullptr, nw, grouped32 ? nullptr : c->wcost, 0};
    int grid = nb8, tpb = CSG_TPB;
    int gw = 1, capg = 0;
    int split = 1;
    if (walk_order_params(c, precision, tape.rec != nullptr, nw, &gw, &capg, &split)) {
        if (!(pre && c->order_gw == gw && c->order_capg == capg && c->order_nw == nw &&
              c->order_split == split)) {
            const int *oc = order_costs(c, nw);
            k_wave_order<<<WO_XCDS, 1024, 0, st>>>(oc, nw, gw, split, capg, c->worder);
        }
        wo.order = c->worder;
        wo.slots = WO_XCDS * capg * gw;
        grid = WO_XCDS * capg;
        tpb = 64 * gw;
    }
    if (!grouped32) c->wcost_n = nw;
    if (precision == COSMOSIM_PREC_FP32_LIST) {
        // wave-shared interaction lists over groups of 64 bodies consecutive in
        // Peano-Hilbert order, gw independent wavefronts per workgroup
        const bool hil = hilbert && c->widx_valid;
        const int *sidx_l = hil ? c->widx : c->idx;
        const int wpb = wo.order ? gw : c->list_block_waves;
        if (!wo.order) grid = blocks_for(nw, wpb);
        const bool selfx = !((float)eps2 >= 1e-12f);
        // the total alone comes from the list lengths where every own leaf is listed
        const bool own_listed = th2 < (1.0 / 3.0) * (1.0 - 1e-5);
        const int cnt = lane_counts ? 2 : (count_total ? (own_listed ? 1 : 2) : 0);
        const int *spos = hil ? c->hpos : nullptr;
        switch (wpb) {
        case 2: launch_walk32l<2>(c, cnt, selfx, sidx_l, spos, G, eps2, th2, grid, wo); break;
        case 4: launch_walk32l<4>(c, cnt, selfx, sidx_l, spos, G, eps2, th2, grid, wo); break;
        default: launch_walk32l<1>(c, cnt, selfx, sidx_l, spos, G, eps2, th2, grid, wo); break;
        }
        CSG_LAUNCHED();
        c->counts_valid = lane_counts ? 1 : 0;
        return COSMOSIM_OK;
    }
    if (wave_mask32) {
        // wave-mask traversal (also the taped FP32 walk: the tape is per wavefront)
        const bool off32 = (size_t)c->node_cap * sizeof(WNode32) < ((size_t)1 << 32);
        const bool taping = tape.rec != nullptr;
        const bool hil = hilbert && c->widx_valid && nw > c->resident_waves;
        const int *sidx = hil ? c->widx : c->idx;
        if (taping) c->tape_wmap = hil ? 1 : 0;
        const int v = (lane_counts ? 4 : 0) | (taping ? 2 : 0) | (off32 ? 1 : 0);
        if (v < 2 && !count_total) {
            // softened: the own leaf's offset is exactly 0 and its factor
            // m / eps^3 is finite in FP32 (eps^2 >= 1e-12 keeps it below
            // m * 1e18), so it adds exactly 0 and needs no exclusion
            if (off32 && (float)eps2 >= 1e-12f)
                launch_walk32w<false, false, true, false, false>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb);
            else if (off32)
                launch_walk32w<false, false, true, false, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb);
            else
                launch_walk32w<false, false, false, false, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb);
        } else {
            switch (v) {
            case 0: launch_walk32w<false, false, false, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 1: launch_walk32w<false, false, true, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 2: launch_walk32w<false, true, false, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 3: launch_walk32w<false, true, true, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 4: launch_walk32w<true, false, false, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 5: launch_walk32w<true, false, true, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 6: launch_walk32w<true, true, false, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            case 7: launch_walk32w<true, true, true, true, true>(c, sidx, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
            }
        }
        CSG_LAUNCHED();
        c->counts_valid = lane_counts ? 1 : 0;
        return COSMOSIM_OK;
    }
    c->counts_valid = 1;
    if (tape.rec) c->tape_wmap = 0;
    // the untaped FP64 walks take their bodies in Peano-Hilbert order (results do not
    // depend on the grouping); a taped walk keeps sorted order
    const int *sidx64 = (hilbert && !tape.rec && c->widx_valid) ? c->widx : c->idx;
    const int *sidx64w = nw > c->resident_waves ? sidx64 : c->idx;   // the wave-mask walk
    // FP64: the wave-mask walk (also for tapes, which are per wavefront) when the
    // records span < 4 GiB, else the grouped walk
    const int g64 = c->walk_group >= 0 ? c->walk_group
                                       : (precision == COSMOSIM_PREC_FP64_STRICT ? 8 : 0);
    if (precision != COSMOSIM_PREC_FP32 && (g64 == 0 || tape.rec) &&
        (size_t)c->node_cap * sizeof(WNode) < ((size_t)1 << 32)) {
        const bool taping = tape.rec != nullptr;
        if (precision == COSMOSIM_PREC_FP64_STRICT) {
            if (taping) launch_walk64w<WALK_FP64_STRICT, true>(c, c->idx, G, eps2, th2, grid, nb8, tape, wo, tpb);
            else launch_walk64w<WALK_FP64_STRICT, false>(c, sidx64w, G, eps2, th2, grid, nb8, tape, wo, tpb);
        } else {
            if (taping) launch_walk64w<WALK_FP64, true>(c, c->idx, G, eps2, th2, grid, nb8, tape, wo, tpb);
            else launch_walk64w<WALK_FP64, false>(c, sidx64w, G, eps2, th2, grid, nb8, tape, wo, tpb);
        }
        CSG_LAUNCHED();
        return COSMOSIM_OK;
    }
    // a taped walk runs with whole-wave traversal (the tape is per wavefront)
    const int gs = tape.rec ? 64 : (precision == COSMOSIM_PREC_FP32 ? c->walk_group32 : g64);
    switch (gs) {
    case 8: launch_walk_gs<8>(c, precision, sidx64, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
    case 16: launch_walk_gs<16>(c, precision, sidx64, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
    case 32: launch_walk_gs<32>(c, precision, sidx64, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
    default: launch_walk_gs<64>(c, precision, sidx64, G, eps2, th2, grid, nb8, tape, wo, tpb); break;
    }
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

int csg_kick_drift(csg_ctx *c, double dt) {
    CSG_USE(c);
    timer_start(c);
    int s = launch_kick_drift(c, dt, 1);
    if (s) return s;
    timer_stop(c, PH_KDK);
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_kick(csg_ctx *c, double dt) {
    CSG_USE(c);
    timer_start(c);
    int s = launch_kick(c, dt);
    if (s) return s;
    timer_stop(c, PH_KDK);
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

// After a run of builds: fails when the last build or any build since the
// previous check overflowed or stalled (the walks after such a build do
// nothing, so the accelerations the run went on with are stale).
static int check_overflow(csg_ctx *c) {
    int s = read_params(c);
    if (s) return s;
    collect_timers(c);
    const int pending = (c->hbp->overflow || c->hbp->sort_stall || c->hbp->order_bad) ? 1 : 0;
    const bool earlier = c->hbp->fail_count != c->fail_seen;
    c->fail_seen = c->hbp->fail_count + pending;
    if (pending || earlier) {
        c->tree_valid = 0;
        c->counts_valid = 0;
        if (c->hbp->sort_stall) {
            fprintf(stderr, "cosmosim: body sort stalled waiting for an earlier tile\n");
            return COSMOSIM_E_HIP;
        }
        if (c->hbp->order_bad) {
            fprintf(stderr, "cosmosim: Peano-Hilbert wave order is not a permutation\n");
            return COSMOSIM_E_HIP;
        }
        return COSMOSIM_E_CAPACITY;
    }
    return COSMOSIM_OK;
}

int csg_build(csg_ctx *c) {
    CSG_USE(c);
    timer_start(c);
    int s = launch_build(c);
    if (s) return s;
    timer_stop(c, PH_BUILD);
    return check_overflow(c);
}

int csg_walk(csg_ctx *c, int precision, double G, double softening, double theta) {
    CSG_USE(c);
    if (!c->tree_valid) return COSMOSIM_E_STATE;
    if (!prec_valid(precision)) return COSMOSIM_E_ARG;
    int s0 = ensure_nodes64(c, precision);
    if (s0) return s0;
    timer_start(c);
    int s = launch_walk(c, precision, G, softening, theta);
    if (s) return s;
    timer_stop(c, PH_WALK);
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

// total: keep the interaction total of the last walk (the statistics report it);
// the adjoint's internal one-substep calls skip it.
static int substeps(csg_ctx *c, int k, double dt, double G, double softening, double theta,
                    int precision, bool total = true);

int csg_substeps(csg_ctx *c, int k, double dt, double G, double softening, double theta,
                 int precision) {
    CSG_USE(c);
    return substeps(c, k, dt, G, softening, theta, precision);
}

static int substeps(csg_ctx *c, int k, double dt, double G, double softening, double theta,
                    int precision, bool total) {
    c->theta_hint = theta;
    if (!prec_valid(precision) || k < 0) return COSMOSIM_E_ARG;
    if (k == 0) return COSMOSIM_OK;
    int s, bbox = 0;
    timer_start(c);
    if ((s = launch_kick_drift(c, dt, 1, &bbox, precision))) return s;
    timer_stop(c, PH_KDK);
    for (int sub = 0; sub < k; ++sub) {
        timer_start(c);
        if ((s = launch_build(c, !prec_rec32(precision), bbox,
                              hilbert_order_for(c, precision))))
            return s;
        timer_stop(c, PH_BUILD);
        timer_start(c);
        if ((s = launch_walk(c, precision, G, softening, theta, TapeOut{}, false,
                             total && sub + 1 == k)))
            return s;
        timer_stop(c, PH_WALK);
        timer_start(c);
        // end-of-substep half kick fused with the next substep's kick + drift
        // (which also bounds the drifted positions for the next build)
        s = (sub + 1 < k) ? launch_kick_drift(c, dt, 2, &bbox, precision) : launch_kick(c, dt);
        if (s) return s;
        timer_stop(c, PH_KDK);
    }
    return check_overflow(c);
}

int csg_export_pool(csg_ctx *c, OctreeNode *pool, int capacity, int *pool_size) {
    CSG_USE(c);
    if (!pool || !pool_size) return COSMOSIM_E_ARG;
    if (!c->tree_valid) return COSMOSIM_E_STATE;
    int s = ensure_nodes64(c, COSMOSIM_PREC_FP64);
    if (s) return s;
    if ((s = read_params(c))) return s;
    const int nn = c->hbp->n_nodes;
    if (c->hbp->overflow || nn > capacity) return COSMOSIM_E_CAPACITY;
    if ((size_t)nn > c->pool_dev_cap) {
        dfree(c->pool_dev);
        c->pool_dev_cap = 0;
        if (!dalloc(c->pool_dev, (size_t)nn)) {
            (void)hipGetLastError();
            return COSMOSIM_E_ALLOC;
        }
        c->pool_dev_cap = (size_t)nn;
    }
    const int g = blocks_for(nn, CSG_TPB);
    k_export_nodes<<<g, CSG_TPB, 0, c->stream>>>(c->nodes, c->key, c->ext, c->bp, c->pool_dev);
    k_export_link<<<g, CSG_TPB, 0, c->stream>>>(c->nodes, c->key, c->ext, c->lcp, c->nbase, c->bp,
                                                c->pool_dev);
    CSG_LAUNCHED();
    CSG_CHECK(hipMemcpyAsync(pool, c->pool_dev, sizeof(OctreeNode) * (size_t)nn,
                             hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    *pool_size = nn;
    return COSMOSIM_OK;
}

int csg_walk_pool(csg_ctx *c, const OctreeNode *pool, int pool_size, int root, double G,
                  double softening, double theta) {
    CSG_USE(c);
    if (!pool || pool_size <= 0 || root < 0 || root >= pool_size) return COSMOSIM_E_ARG;
    if ((size_t)pool_size > c->pool_dev_cap) {
        dfree(c->pool_dev);
        c->pool_dev_cap = 0;
        if (!dalloc(c->pool_dev, (size_t)pool_size)) {
            (void)hipGetLastError();
            return COSMOSIM_E_ALLOC;
        }
        c->pool_dev_cap = (size_t)pool_size;
    }
    CSG_CHECK(hipMemcpyAsync(c->pool_dev, pool, sizeof(OctreeNode) * (size_t)pool_size,
                             hipMemcpyHostToDevice, c->stream));
    k_set_i32<<<1, 1, 0, c->stream>>>(c->err_dev, 0);
    if (c->n > 0) {
        k_walk_pool<<<blocks_for(c->n, CSG_TPB), CSG_TPB, 0, c->stream>>>(
            c->pool_dev, pool_size, root, c->x, c->y, c->z, c->m, c->n, c->ax, c->ay, c->az,
            c->icount, G, softening * softening, theta, c->err_dev);
    }
    CSG_LAUNCHED();
    c->counts_valid = 1;
    int err = 0;
    CSG_CHECK(hipMemcpyAsync(&err, c->err_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    if (err & 2) return COSMOSIM_E_ARG;
    if (err & 1) return COSMOSIM_E_CAPACITY;
    return COSMOSIM_OK;
}

int csg_direct_forces(csg_ctx *c, const int *idx, int k, double G, double softening,
                      double *acc) {
    CSG_USE(c);
    if (k < 0 || (k > 0 && (!idx || !acc))) return COSMOSIM_E_ARG;
    if (k == 0 || c->n == 0) {
        for (int t = 0; t < 3 * k; ++t) acc[t] = 0.0;
        return COSMOSIM_OK;
    }
    for (int t = 0; t < k; ++t)
        if (idx[t] < 0 || idx[t] >= c->n) return COSMOSIM_E_ARG;
    const int n = c->n;
    const int groups = blocks_for(k, 64);
    int S = 1;
    if (n > 4096) {
        S = blocks_for(4096, groups);
        if (S > 512) S = 512;
        if (S < 1) S = 1;
    }
    int chunk = blocks_for(n, S);
    S = blocks_for(n, chunk);
    const size_t part_bytes = sizeof(double) * 3 * (size_t)k * S;
    const size_t need = part_bytes + sizeof(double) * 3 * (size_t)k + sizeof(int) * (size_t)k + 256;
    int s = ensure_diag(c, need);
    if (s) return s;
    double *part = (double *)c->diag;
    double *out = part + 3 * (size_t)k * S;
    int *didx = (int *)(out + 3 * (size_t)k);
    CSG_CHECK(hipMemcpyAsync(didx, idx, sizeof(int) * k, hipMemcpyHostToDevice, c->stream));
    k_direct<<<dim3(groups, S), 64, 0, c->stream>>>(didx, k, c->x, c->y, c->z, c->m, n, chunk, G,
                                                    softening * softening, part);
    k_direct_reduce<<<blocks_for(k, CSG_TPB), CSG_TPB, 0, c->stream>>>(part, k, S, out);
    CSG_LAUNCHED();
    CSG_CHECK(hipMemcpyAsync(acc, out, sizeof(double) * 3 * k, hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_energy(csg_ctx *c, double G, double softening, double *kinetic, double *potential,
               double momentum[3]) {
    CSG_USE(c);
    const int n = c->n;
    double res[5] = {0, 0, 0, 0, 0};
    if (n > 0) {
        const int nbk = blocks_for(n, CSG_TPB);
        int S = blocks_for(4096, nbk);
        if (S < 1) S = 1;
        if (S > 1024) S = 1024;
        int chunk = blocks_for(n, S);
        S = blocks_for(n, chunk);
        const size_t pe_parts = (size_t)nbk * S;
        const size_t need = sizeof(double) * (4 * (size_t)nbk + pe_parts + 8);
        int s = ensure_diag(c, need);
        if (s) return s;
        double *kp = (double *)c->diag;
        double *pp = kp + 4 * (size_t)nbk;
        double *out = pp + pe_parts;
        k_kinetic<<<nbk, CSG_TPB, 0, c->stream>>>(c->vx, c->vy, c->vz, c->m, n, kp);
        k_potential<<<dim3(nbk, S), CSG_TPB, 0, c->stream>>>(c->x, c->y, c->z, c->m, n, chunk, G,
                                                             softening * softening, pp);
        for (int q = 0; q < 4; ++q)
            k_sum<<<1, CSG_TPB, 0, c->stream>>>(kp, nbk, 4, q, out + q);
        k_sum<<<1, CSG_TPB, 0, c->stream>>>(pp, (int)pe_parts, 1, 0, out + 4);
        CSG_LAUNCHED();
        CSG_CHECK(hipMemcpyAsync(res, out, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
        CSG_CHECK(hipStreamSynchronize(c->stream));
    }
    if (kinetic) *kinetic = res[0];
    if (momentum) { momentum[0] = res[1]; momentum[1] = res[2]; momentum[2] = res[3]; }
    if (potential) *potential = -0.5 * res[4];
    return COSMOSIM_OK;
}

int csg_interaction_counts(csg_ctx *c, int *counts, int n) {
    CSG_USE(c);
    if (n != c->n || (n > 0 && !counts)) return COSMOSIM_E_ARG;
    if (n == 0) return COSMOSIM_OK;
    if (!c->counts_valid) return COSMOSIM_E_STATE;
    CSG_CHECK(hipMemcpyAsync(counts, c->icount, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

int csg_get_stats(csg_ctx *c, csg_stats *o) {
    CSG_USE(c);
    if (!o) return COSMOSIM_E_ARG;
    if (c->tree_valid && c->n > 0) {   // the deepest leaf is found on demand
        k_max_depth<<<blocks_for(c->cap, CSG_TILE), CSG_TPB, 0, c->stream>>>(c->lcp, c->bp);
        CSG_LAUNCHED();
    }
    int s = read_params(c);
    if (s) return s;
    unsigned long long it = 0;
    int fall = 0;
    CSG_CHECK(hipMemcpyAsync(&it, c->itotal, sizeof(it), hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipMemcpyAsync(&fall, c->wl_fall, sizeof(fall), hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    collect_timers(c);
    memset(o, 0, sizeof(*o));
    o->n = c->n;
    o->n_live = c->hbp->n_live;
    o->n_nodes = c->hbp->n_nodes;
    o->max_depth = c->hbp->max_depth;
    o->collision_runs = c->hbp->n_runs;
    o->overflow = c->hbp->overflow;
    o->interactions = (long long)it;
    for (int a = 0; a < 4; ++a) o->root[a] = c->hbp->root[a];
    o->ms_build = c->ms_build;
    o->ms_walk = c->ms_walk;
    o->ms_kdk = c->ms_kdk;
    o->timed_walks = c->timed_walks;
    o->ms_neighbors = c->ms_nb;
    o->ms_sph = c->ms_sph;
    o->neighbor_queries = c->nb_queries;
    o->list_fallbacks = fall;
    return COSMOSIM_OK;
}

int csg_set_timing(csg_ctx *c, int enabled) {
    CSG_USE(c);
    if (enabled && !c->ev) {
        c->ev = (hipEvent_t *)calloc(2 * CSG_EV_MAX, sizeof(hipEvent_t));
        if (!c->ev) return COSMOSIM_E_ALLOC;
        for (int k = 0; k < 2 * CSG_EV_MAX; ++k) CSG_CHECK(hipEventCreate(&c->ev[k]));
    }
    CSG_CHECK(hipStreamSynchronize(c->stream));
    c->timing = enabled ? 1 : 0;
    c->ev_count = 0;
    c->ev_open = -1;
    c->ms_build = c->ms_walk = c->ms_kdk = c->ms_nb = c->ms_sph = 0.0;
    c->timed_walks = 0;
    c->nb_queries = 0;
    clear_words(c->wl_fall, 1, c->stream);
    return COSMOSIM_OK;
}

int csg_debug_fail_build(csg_ctx *c, int nth) {
    CSG_USE(c);
    if (nth < 0) return COSMOSIM_E_ARG;
    c->debug_fail_build = nth;
    return COSMOSIM_OK;
}

int csg_debug_orders(csg_ctx *c, int *sorted_idx, int *wave_idx, unsigned long long *keys, int n,
                     int *n_live) {
    CSG_USE(c);
    if (!n_live || n < 0) return COSMOSIM_E_ARG;
    if (!c->tree_valid) return COSMOSIM_E_STATE;
    int s = read_params(c);
    if (s) return s;
    const int nl = c->hbp->n_live;
    *n_live = nl;
    if (n < nl) return COSMOSIM_E_CAPACITY;
    if (nl == 0) return COSMOSIM_OK;
    if (sorted_idx)
        CSG_CHECK(hipMemcpyAsync(sorted_idx, c->idx, sizeof(int) * nl, hipMemcpyDeviceToHost, c->stream));
    if (wave_idx)
        CSG_CHECK(hipMemcpyAsync(wave_idx, c->widx_valid ? c->widx : c->idx, sizeof(int) * nl,
                                 hipMemcpyDeviceToHost, c->stream));
    if (keys)
        CSG_CHECK(hipMemcpyAsync(keys, c->key, sizeof(uint64_t) * nl, hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

}  // extern "C"

// ================================================ allocation-free force call
// cosmosim_ext_forces_noalloc: the tree build and walk of
// cosmosim_ext_compute_forces on caller-supplied device arrays, with every
// other buffer carved out of one file-scope device array instead of hipMalloc:
// keys and sort buffers, collision words, sorted copies, prefix sums, node
// records, dispatch and counter words.  The array holds the regions of at most
// CSG_NOALLOC_MAX_BODIES bodies; each GPU the library runs on keeps its own
// copy until the process ends, and every call on that GPU reuses it.
#define CSG_NOALLOC_MAX_BODIES 16384
enum {
    NA_BBOX, NA_KEY, NA_KEY2, NA_EXT, NA_EXT_TMP, NA_IDX, NA_IDX2, NA_WIDX, NA_XS, NA_YS, NA_ZS,
    NA_MS,
    NA_LCP, NA_LMIN1, NA_LMIN2, NA_CNT, NA_NBASE, NA_SCAN, NA_PREF,
    NA_DDPART, NA_TPREF,
    NA_NODES, NA_NODES32, NA_NLAST, NA_ICOUNT, NA_WCOST, NA_WORDER, NA_XYZM, NA_RSCTL,
    NA_RSSTATUS, NA_BP, NA_ITOTAL, NA_REGIONS
};
struct NoallocLayout {
    size_t off[NA_REGIONS];
    size_t total;
};
// Byte offset of every region for `cap` bodies, each 256-byte aligned; the
// sizes are alloc_buffers' for the buffers launch_build and launch_walk use.
constexpr NoallocLayout noalloc_layout(int cap) {
    NoallocLayout L{};
    const size_t C = (size_t)(cap < 1 ? 1 : cap), C1 = C + 1;
    const size_t nb = (C1 + CSG_TILE - 1) / CSG_TILE;
    const size_t nodes = 8 * C + 4096, waves = C / 64 + 1;
    const size_t bytes[NA_REGIONS] = {
        (size_t)CSG_BBOX_BLOCKS * 8 * sizeof(double),
        C * sizeof(uint64_t), C * sizeof(uint64_t), 3 * C * sizeof(uint64_t),
        3 * C * sizeof(uint64_t), C * sizeof(int), C * sizeof(int), C * sizeof(int),
        C * sizeof(double), C * sizeof(double), C * sizeof(double), C * sizeof(double),
        C1 + 64,
        lmin1_bytes(C), lmin2_bytes(C), C1 * sizeof(int), C1 * sizeof(int), nb * sizeof(int), 4 * C1 * sizeof(DD),
        4 * nb * sizeof(DD), 4 * (size_t)CSG_TPB * nb * sizeof(DD),
        nodes * sizeof(WNode), nodes * sizeof(WNode32), nodes * sizeof(int), C * sizeof(int),
        waves * sizeof(int),
        (2 * waves + 4096) * sizeof(int), C * sizeof(double4), sizeof(RsCtl), 256 * ((C + RS_TILE - 1) / RS_TILE) * sizeof(uint64_t),
        sizeof(BuildParams), sizeof(unsigned long long)};
    size_t o = 0;
    for (int r = 0; r < NA_REGIONS; ++r) {
        L.off[r] = o;
        o += (bytes[r] + 255) / 256 * 256;
    }
    L.total = o;
    return L;
}
#define CSG_NOALLOC_BYTES (noalloc_layout(CSG_NOALLOC_MAX_BODIES).total)
// The host must not use this name as a device address: kernels get the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[CSG_NOALLOC_BYTES];

#define CSG_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[CSG_MAX_DEVICES];

// The scratch array's device address on `dev` (the current device), looked up
// once per device.
static int noalloc_scratch(int dev, unsigned char **out) {
    if (dev < 0 || dev >= CSG_MAX_DEVICES) return COSMOSIM_E_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "cosmosim: hipGetSymbolAddress(scratch) failed: %s\n",
                    hipGetErrorString(e));
            (void)hipGetLastError();
            return COSMOSIM_E_HIP;
        }
        g_scratch_ptr[dev] = (unsigned char *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return COSMOSIM_OK;
}

// status: 0, or the code a build that overflowed (or stalled) would return.
__global__ void k_noalloc_status(const BuildParams *bp, int *status) {
    *status = (bp->sort_stall || bp->order_bad) ? COSMOSIM_E_HIP
                                                : (bp->overflow ? COSMOSIM_E_CAPACITY : 0);
}

extern "C" {

int cosmosim_ext_forces_noalloc(const double *x, const double *y, const double *z,
                                const double *m, int n, double G, double softening,
                                double theta, int precision, double *ax, double *ay,
                                double *az, int *status, void *stream) {
    if (n < 0 || (n > 0 && (!x || !y || !z || !m || !ax || !ay || !az)) || !status ||
        precision < COSMOSIM_PREC_FP64_STRICT || precision > COSMOSIM_PREC_FP32 ||
        !(theta >= 0.0) || !(softening >= 0.0))
        return COSMOSIM_E_ARG;
    if (n > CSG_NOALLOC_MAX_BODIES) {
        fprintf(stderr, "cosmosim: cosmosim_ext_forces_noalloc takes at most %d bodies (%d)\n",
                CSG_NOALLOC_MAX_BODIES, n);
        return COSMOSIM_E_CAPACITY;
    }
    int dev = 0;
    CSG_CHECK(hipGetDevice(&dev));
    unsigned char *base = nullptr;
    int s = noalloc_scratch(dev, &base);
    if (s) return s;
    const NoallocLayout L = noalloc_layout(n);
    if (L.total > CSG_NOALLOC_BYTES) return COSMOSIM_E_CAPACITY;   // not reached: n <= max
    auto at = [&](int r) { return (void *)(base + L.off[r]); };
    // a context on the stack whose buffers are the caller's arrays and the scratch
    csg_ctx c;
    c.device = dev;
    c.stream = (hipStream_t)stream;
    c.cap = n > 0 ? n : 1;
    c.n = n;
    c.node_cap = 8 * c.cap + 4096;
    c.x = (double *)x; c.y = (double *)y; c.z = (double *)z; c.m = (double *)m;
    c.ax = ax; c.ay = ay; c.az = az;
    c.bbox_part = (double *)at(NA_BBOX);
    c.key = (uint64_t *)at(NA_KEY); c.key2 = (uint64_t *)at(NA_KEY2);
    c.ext = (uint64_t *)at(NA_EXT); c.ext_tmp = (uint64_t *)at(NA_EXT_TMP);
    c.idx = (int *)at(NA_IDX); c.idx2 = (int *)at(NA_IDX2); c.widx = (int *)at(NA_WIDX);
    c.xs = (double *)at(NA_XS); c.ys = (double *)at(NA_YS); c.zs = (double *)at(NA_ZS);
    c.ms = (double *)at(NA_MS);
    c.lcp = (signed char *)at(NA_LCP);
    c.lmin1 = (signed char *)at(NA_LMIN1); c.lmin2 = (signed char *)at(NA_LMIN2);
    c.cnt = (int *)at(NA_CNT); c.nbase = (int *)at(NA_NBASE);
    c.scan_part = (int *)at(NA_SCAN); c.pref = (DD *)at(NA_PREF); c.dd_part = (DD *)at(NA_DDPART);
    c.tpref = (DD *)at(NA_TPREF); c.nodes = (WNode *)at(NA_NODES);
    c.nodes32 = (WNode32 *)at(NA_NODES32); c.nlast = (int *)at(NA_NLAST);
    c.icount = (int *)at(NA_ICOUNT);
    c.wcost = (int *)at(NA_WCOST); c.worder = (int *)at(NA_WORDER);
    c.xyzm = (double4 *)at(NA_XYZM);
    c.rs_ctl = (RsCtl *)at(NA_RSCTL); c.rs_status = (uint64_t *)at(NA_RSSTATUS);
    c.bp = (BuildParams *)at(NA_BP); c.itotal = (unsigned long long *)at(NA_ITOTAL);
    // the array keeps the last call's contents: start the words the build reads
    // before writing them from zero (fill kernels, no memset nodes)
    clear_words((int *)c.bp, sizeof(BuildParams) / sizeof(int), c.stream);
    clear_words(c.rs_status, 256 * (size_t)blocks_for(c.cap, RS_TILE), c.stream);
    clear_words(c.idx, (size_t)c.cap, c.stream);
    clear_words(c.idx2, (size_t)c.cap, c.stream);
    clear_words(c.widx, (size_t)c.cap, c.stream);
    CSG_LAUNCHED();
    c.theta_hint = theta;
    if ((s = launch_build(&c, precision != COSMOSIM_PREC_FP32, 0, hilbert_order_for(&c, precision))))
        return s;
    if ((s = launch_walk(&c, precision, G, softening, theta))) return s;
    k_noalloc_status<<<1, 1, 0, c.stream>>>(c.bp, status);
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

int csg_synchronize(csg_ctx *c) {
    CSG_USE(c);
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

void *csg_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 1;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    memset(p, 0, bytes);
    return p;
}

void csg_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"

#include "csg_adjoint.inl"
#include "csg_neighbors.inl"
#include "csg_sph.inl"
#include "csg_snapshot.inl"
#include "csg_quasar.inl"
