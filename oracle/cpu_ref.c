/*
 * cpu_ref.c - CPU oracle for the gravity hot path.  TEST INFRASTRUCTURE ONLY:
 * it is linked by nothing in the product; tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py load it (through oracle/cpu_ref.py) as the
 * checker and as the timed host baseline.
 *
 * It restates, in its own code, the semantics of
 *   src/octree.c:5-234     (tree build by sequential insertion, monopole walk)
 *   src/integrator.c:4-57  (KDK leapfrog, without the SPH hooks)
 *   src/octree.c:236-318   (sphere neighbour query)
 *   src/sph.c:10-231       (SPH density, forces and cooling over those queries)
 *   CosmosimSubsystem.cpp:12-43, src/renderer.c:348-364 (float snapshots)
 *   src/quasar.c:11-529    (quasar step and compaction, explicit random state)
 *   tests/test_physics.c:34-71 (direct-sum and energy judges)
 * and the core's own FP32 walk on wave-shared interaction lists (cref_list_walk,
 * which has no counterpart in the reference).
 * Floating-point expressions are kept in the reference's evaluation order so
 * that, compiled with -O2 -ffp-contract=off on x86-64, results are bit-identical
 * to the reference (checked by tests/test_oracle.py against tests/golden/,
 * which tests/golden/gen_golden.py wrote from the compiled reference).
 *
 * Differences in structure, not in results: insertion and the walk are
 * iterative (explicit loops / stack) instead of recursive, and the walk can
 * report the number of interactions per body.
 */
#include <float.h>
#include <stdint.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "body.h"
#include "octree.h"

#define CREF_DEPTH_LIMIT 64 /* src/octree.c:5 */

static inline int cref_live(const Body *b) { return !(b->mass <= 0.0); }

/* New node with no mass, no body and no children (src/octree.c:7-28). */
static int cref_node(OctreeNode *pool, int *used, double gx, double gy, double gz, double half)
{
    int id = (*used)++;
    OctreeNode *nd = &pool[id];
    memset(nd, 0, sizeof(*nd));
    nd->center_x = gx;
    nd->center_y = gy;
    nd->center_z = gz;
    nd->half_size = half;
    nd->body_index = -1;
    for (int c = 0; c < 8; ++c)
        nd->children[c] = -1;
    return id;
}

/* Octant code: bit0 east (x >= centre), bit1 south (y < centre), bit2 bottom
 * (z < centre) (src/octree.c:30-40). */
static inline int cref_octant(const OctreeNode *nd, double x, double y, double z)
{
    return (x >= nd->center_x ? 1 : 0) | (y < nd->center_y ? 2 : 0) | (z < nd->center_z ? 4 : 0);
}

/* Child cell of octant q: centre shifted by half the parent half-size along each
 * axis (+ for east, - for south/bottom), half-size halved (src/octree.c:42-50,
 * 87-90, 105-108). */
static int cref_child(OctreeNode *pool, int *used, int parent, int q)
{
    const OctreeNode *p = &pool[parent];
    double h = p->half_size;
    double sx = (q & 1) ? 0.5 : -0.5;
    double sy = (q & 2) ? -0.5 : 0.5;
    double sz = (q & 4) ? -0.5 : 0.5;
    double gx = p->center_x + sx * h;
    double gy = p->center_y + sy * h;
    double gz = p->center_z + sz * h;
    return cref_node(pool, used, gx, gy, gz, h * 0.5);
}

/* Running-average mass/COM update in the reference's order (src/octree.c:96-100). */
static inline void cref_absorb(OctreeNode *nd, double x, double y, double z, double m)
{
    double mm = nd->total_mass + m;
    nd->cx = (nd->cx * nd->total_mass + x * m) / mm;
    nd->cy = (nd->cy * nd->total_mass + y * m) / mm;
    nd->cz = (nd->cz * nd->total_mass + z * m) / mm;
    nd->total_mass = mm;
}

/* Insert one live body, walking down from the root (src/octree.c:52-112). */
static void cref_insert(OctreeNode *pool, int *used, const Body *bodies, int b)
{
    const double x = bodies[b].x, y = bodies[b].y, z = bodies[b].z, m = bodies[b].mass;
    int v = 0;
    for (int depth = 0;; ++depth) {
        OctreeNode *nd = &pool[v];
        if (nd->total_mass == 0.0 && nd->body_index == -1) {
            /* empty cell: the body becomes its leaf */
            nd->body_index = b;
            nd->cx = x;
            nd->cy = y;
            nd->cz = z;
            nd->total_mass = m;
            return;
        }
        if (depth >= CREF_DEPTH_LIMIT) {
            /* depth limit: merge into this leaf, first body index kept */
            cref_absorb(nd, x, y, z, m);
            return;
        }
        if (nd->body_index >= 0) {
            /* occupied leaf: push the resident body one level down */
            int r = nd->body_index;
            nd->body_index = -1;
            int qr = cref_octant(nd, bodies[r].x, bodies[r].y, bodies[r].z);
            int c = cref_child(pool, used, v, qr);
            nd = &pool[v];
            nd->children[qr] = c;
            OctreeNode *cn = &pool[c];
            cn->body_index = r;
            cn->cx = bodies[r].x;
            cn->cy = bodies[r].y;
            cn->cz = bodies[r].z;
            cn->total_mass = bodies[r].mass;
        }
        cref_absorb(nd, x, y, z, m);
        int q = cref_octant(nd, x, y, z);
        if (nd->children[q] < 0) {
            int c = cref_child(pool, used, v, q);
            pool[v].children[q] = c;
        }
        v = pool[v].children[q];
    }
}

/* Root cube over live bodies and sequential insertion (src/octree.c:114-169). */
void cref_build(OctreeNode *pool, int *pool_size, const Body *bodies, int n)
{
    int used = 0;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    int live = 0;
    for (int i = 0; i < n; ++i) {
        if (!cref_live(&bodies[i]))
            continue;
        ++live;
        const double p[3] = {bodies[i].x, bodies[i].y, bodies[i].z};
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a])
                lo[a] = p[a];
            if (p[a] > hi[a])
                hi[a] = p[a];
        }
    }
    if (live == 0) {
        cref_node(pool, &used, 0.0, 0.0, 0.0, 1.0);
        *pool_size = used;
        return;
    }
    double ext = hi[0] - lo[0];
    if (hi[1] - lo[1] > ext)
        ext = hi[1] - lo[1];
    if (hi[2] - lo[2] > ext)
        ext = hi[2] - lo[2];
    double half = ext * 0.5 * 1.01;
    if (half < 1e-10)
        half = 1.0;
    cref_node(pool, &used, (lo[0] + hi[0]) * 0.5, (lo[1] + hi[1]) * 0.5, (lo[2] + hi[2]) * 0.5,
              half);
    for (int i = 0; i < n; ++i)
        if (cref_live(&bodies[i]))
            cref_insert(pool, &used, bodies, i);
    *pool_size = used;
}

/* Monopole term of a node or body on body i, added in the reference's order
 * (src/octree.c:192-197, 204-209). */
static inline void cref_term(Body *bi, double G, double M, double dx, double dy, double dz,
                             double d2, double eps2)
{
    double r2 = d2 + eps2;
    double w = 1.0 / (r2 * sqrt(r2));
    bi->ax += G * M * dx * w;
    bi->ay += G * M * dy * w;
    bi->az += G * M * dz * w;
}

/* Depth-first walk for one body with an explicit stack; children in slot order
 * 0..7 (src/octree.c:171-218).  Returns the number of interactions. */
static long cref_walk_one(const OctreeNode *pool, int root, Body *bodies, int i, double G,
                          double eps2, double theta)
{
    /* pending[] holds nodes still to visit, pushed in reverse slot order so
     * they pop in slot order. */
    int pending[8 * 70 + 8];
    int top = 0;
    long count = 0;
    Body *bi = &bodies[i];
    const double th2 = theta * theta;
    pending[top++] = root;
    while (top > 0) {
        const OctreeNode *nd = &pool[pending[--top]];
        if (nd->total_mass == 0.0)
            continue;
        double dx = nd->cx - bi->x;
        double dy = nd->cy - bi->y;
        double dz = nd->cz - bi->z;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (nd->body_index >= 0) {
            if (nd->body_index != i) {
                cref_term(bi, G, nd->total_mass, dx, dy, dz, d2, eps2);
                ++count;
            }
            continue;
        }
        double s = 2.0 * nd->half_size;
        if (d2 > 0.0 && (s * s) / d2 < th2) {
            cref_term(bi, G, nd->total_mass, dx, dy, dz, d2, eps2);
            ++count;
            continue;
        }
        for (int c = 7; c >= 0; --c)
            if (nd->children[c] >= 0)
                pending[top++] = nd->children[c];
    }
    return count;
}

/* Zero a for every body, then walk every live body (src/octree.c:220-234).
 * counts (may be NULL) receives per-body interaction counts. */
void cref_forces(const OctreeNode *pool, int root, Body *bodies, int n, double G,
                 double softening, double theta, int *counts)
{
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        bodies[i].ax = 0.0;
        bodies[i].ay = 0.0;
        bodies[i].az = 0.0;
        long c = 0;
        if (cref_live(&bodies[i]))
            c = cref_walk_one(pool, root, bodies, i, G, eps2, theta);
        if (counts)
            counts[i] = (int)c;
    }
}

/* ---------------------------------------------------------------- list walk
 * Restatement of the core's FP32 walk on wave-shared interaction lists
 * (k_walk32l; no counterpart in the reference, whose walk is per body,
 * src/octree.c:171-218).  The bodies are taken in groups of 64 consecutive
 * entries of `order` (the device's wavefront order).  For each group:
 *   - its bounding box in FP64; centre g = 0.5 (lo + hi); half extents
 *     w = fp32(0.5 (hi - lo)) * (1 + 2^-20) in FP32;
 *   - a depth-first walk over the tree (children in slot order) in which a
 *     leaf is always listed and an internal node is listed when
 *     s2 < dmin2, with s2 = fp32((2h)^2 / theta^2) (FP64 product, +inf at
 *     theta 0) and dmin2 = fmaf(tx, tx, fmaf(ty, ty, tz * tz)),
 *     t = fmaxf(|fp32(C - g)| - w, 0) per axis, all FP32 (the device's
 *     arithmetic, so the listed set is the device's); otherwise it is opened;
 *   - every body of the group sums the whole list in FP64 (exact offsets),
 *     skipping a leaf whose body_index is its own.
 * acc (n x 3) receives G * sum for live bodies (0 for others), counts (may be
 * NULL) the listed entries each body summed.  checks (may be NULL) receives
 * [0] the largest |listed mass - live mass| / live mass over the groups (every
 * body must be covered exactly once) and [1] the number of listed internal
 * nodes that fail the reference criterion (2h)^2 / d^2 < theta^2 (FP64) for
 * some body of their group by more than 1e-5 relative. */
void cref_list_walk(const OctreeNode *pool, int root, const Body *bodies, int n,
                    const int *order, int n_order, double G, double softening, double theta,
                    double *acc, int *counts, double *checks)
{
    const double eps2 = softening * softening;
    const double inv_th2 = theta > 0.0 ? 1.0 / (theta * theta) : HUGE_VAL;
    double live_mass = 0.0;
    for (int i = 0; i < n; ++i)
        if (cref_live(&bodies[i]))
            live_mass += bodies[i].mass;
    memset(acc, 0, sizeof(double) * 3 * (size_t)n);
    if (counts)
        memset(counts, 0, sizeof(int) * (size_t)n);
    double worst_mass = 0.0;
    long bad = 0;
    const int ngroups = (n_order + 63) / 64;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : worst_mass) reduction(+ : bad)
    for (int g = 0; g < ngroups; ++g) {
        const int b0 = 64 * g, b1 = b0 + 64 < n_order ? b0 + 64 : n_order;
        double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int j = b0; j < b1; ++j) {
            const Body *b = &bodies[order[j]];
            const double p[3] = {b->x, b->y, b->z};
            for (int a = 0; a < 3; ++a) {
                if (p[a] < lo[a]) lo[a] = p[a];
                if (p[a] > hi[a]) hi[a] = p[a];
            }
        }
        double gc[3];
        float w[3];
        for (int a = 0; a < 3; ++a) {
            gc[a] = 0.5 * (lo[a] + hi[a]);
            w[a] = (float)(0.5 * (hi[a] - lo[a])) * (1.0f + 0x1p-20f);
        }
        /* the list: node ids, by an explicit-stack depth-first walk */
        int cap = 1024, len = 0;
        int *list = malloc(sizeof(int) * (size_t)cap);
        int stack[8 * 70 + 8];
        int top = 0;
        stack[top++] = root;
        double listed = 0.0;
        while (top > 0) {
            const int id = stack[--top];
            const OctreeNode *nd = &pool[id];
            int take = nd->body_index >= 0;
            if (!take) {
                const float r[3] = {(float)(nd->cx - gc[0]), (float)(nd->cy - gc[1]),
                                    (float)(nd->cz - gc[2])};
                float t[3];
                for (int a = 0; a < 3; ++a) {
                    const float u = fabsf(r[a]) - w[a];
                    t[a] = u > 0.0f ? u : 0.0f;
                }
                const float dm2 = fmaf(t[0], t[0], fmaf(t[1], t[1], t[2] * t[2]));
                const double s = 2.0 * nd->half_size;
                const float s2 = (float)(s * s * inv_th2);
                take = s2 < dm2;
                if (take) {
                    for (int j = b0; j < b1; ++j) {
                        const Body *b = &bodies[order[j]];
                        const double dx = nd->cx - b->x, dy = nd->cy - b->y, dz = nd->cz - b->z;
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (!(s * s < theta * theta * d2 * (1.0 + 1e-5)))
                            ++bad;
                    }
                }
            }
            if (take) {
                if (len == cap) {
                    cap *= 2;
                    list = realloc(list, sizeof(int) * (size_t)cap);
                }
                list[len++] = id;
                listed += nd->total_mass;
                continue;
            }
            for (int c = 7; c >= 0; --c)
                if (nd->children[c] >= 0)
                    stack[top++] = nd->children[c];
        }
        const double dm = fabs(listed - live_mass) / live_mass;
        if (dm > worst_mass)
            worst_mass = dm;
        for (int j = b0; j < b1; ++j) {
            const int i = order[j];
            const Body *b = &bodies[i];
            double sx = 0.0, sy = 0.0, sz = 0.0;
            int cnt = 0;
            for (int k = 0; k < len; ++k) {
                const OctreeNode *nd = &pool[list[k]];
                if (nd->body_index == i)
                    continue;
                const double dx = nd->cx - b->x, dy = nd->cy - b->y, dz = nd->cz - b->z;
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                const double f = nd->total_mass / (r2 * sqrt(r2));
                sx += f * dx;
                sy += f * dy;
                sz += f * dz;
                ++cnt;
            }
            acc[3 * i] = G * sx;
            acc[3 * i + 1] = G * sy;
            acc[3 * i + 2] = G * sz;
            if (counts)
                counts[i] = cnt;
        }
        free(list);
    }
    if (checks) {
        checks[0] = worst_mass;
        checks[1] = (double)bad;
    }
}

/* Half kick (src/integrator.c:45-51). */
static void cref_half_kick(Body *bodies, int n, double dt)
{
    for (int i = 0; i < n; ++i) {
        Body *b = &bodies[i];
        if (!cref_live(b))
            continue;
        b->vx += 0.5 * b->ax * dt;
        b->vy += 0.5 * b->ay * dt;
        b->vz += 0.5 * b->az * dt;
    }
}

/* Build + walk once (src/integrator.c:4-10). */
void cref_init_acc(Body *bodies, int n, double G, double softening, double theta,
                   OctreeNode *pool)
{
    int used = 0;
    cref_build(pool, &used, bodies, n);
    cref_forces(pool, 0, bodies, n, G, softening, theta, NULL);
}

/* One KDK substep without the SPH hooks (src/integrator.c:12-57). */
void cref_step(Body *bodies, int n, double dt, double G, double softening, double theta,
               OctreeNode *pool)
{
    for (int i = 0; i < n; ++i) {
        Body *b = &bodies[i];
        if (!cref_live(b))
            continue;
        b->vx += 0.5 * b->ax * dt;
        b->vy += 0.5 * b->ay * dt;
        b->vz += 0.5 * b->az * dt;
        b->x += b->vx * dt;
        b->y += b->vy * dt;
        b->z += b->vz * dt;
    }
    int used = 0;
    cref_build(pool, &used, bodies, n);
    cref_forces(pool, 0, bodies, n, G, softening, theta, NULL);
    cref_half_kick(bodies, n, dt);
}

/* Timed pieces for the CPU baseline: the same work as cref_step, split by phase. */
void cref_kick_drift(Body *bodies, int n, double dt)
{
    for (int i = 0; i < n; ++i) {
        Body *b = &bodies[i];
        if (!cref_live(b))
            continue;
        b->vx += 0.5 * b->ax * dt;
        b->vy += 0.5 * b->ay * dt;
        b->vz += 0.5 * b->az * dt;
        b->x += b->vx * dt;
        b->y += b->vy * dt;
        b->z += b->vz * dt;
    }
}

void cref_kick(Body *bodies, int n, double dt) { cref_half_kick(bodies, n, dt); }

/* O(N^2) softened direct sum over all other bodies in index order
 * (tests/test_physics.c:54-71).  Only the k targets idx[] are computed; their
 * accelerations go to acc[k][3]. */
void cref_direct(const Body *bodies, int n, const int *idx, int k, double G, double softening,
                 double *acc)
{
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < k; ++t) {
        const int i = idx[t];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i)
                continue;
            double dx = bodies[j].x - bodies[i].x;
            double dy = bodies[j].y - bodies[i].y;
            double dz = bodies[j].z - bodies[i].z;
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double w = 1.0 / (r2 * sqrt(r2));
            ax += G * bodies[j].mass * dx * w;
            ay += G * bodies[j].mass * dy * w;
            az += G * bodies[j].mass * dz * w;
        }
        acc[3 * t + 0] = ax;
        acc[3 * t + 1] = ay;
        acc[3 * t + 2] = az;
    }
}

/* Softened total energy KE + PE with PE over pairs i<j
 * (tests/test_physics.c:34-52).  out[0] = KE, out[1] = PE. */
void cref_energy(const Body *bodies, int n, double G, double softening, double *out)
{
    double ke = 0.0;
    for (int i = 0; i < n; ++i) {
        const Body *b = &bodies[i];
        ke += 0.5 * b->mass * (b->vx * b->vx + b->vy * b->vy + b->vz * b->vz);
    }
    double pe = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dx = bodies[j].x - bodies[i].x;
            double dy = bodies[j].y - bodies[i].y;
            double dz = bodies[j].z - bodies[i].z;
            double r = sqrt(dx * dx + dy * dy + dz * dz + softening * softening);
            pe -= G * bodies[i].mass * bodies[j].mass / r;
        }
    }
    out[0] = ke;
    out[1] = pe;
}

/* Sphere query: cells whose box misses the sphere are pruned, leaves report
 * live bodies inside the sphere, children in slot order, at most cap results
 * (src/octree.c:236-318). */
static void cref_nb_visit(const OctreeNode *pool, int v, const Body *bodies, double qx, double qy,
                          double qz, double r2, int *buf, int *count, int cap)
{
    const OctreeNode *nd = &pool[v];
    double ex = fmax(0.0, fabs(qx - nd->center_x) - nd->half_size);
    double ey = fmax(0.0, fabs(qy - nd->center_y) - nd->half_size);
    double ez = fmax(0.0, fabs(qz - nd->center_z) - nd->half_size);
    if (ex * ex + ey * ey + ez * ez > r2)
        return;
    if (nd->body_index >= 0) {
        if (*count >= cap)
            return;
        const Body *b = &bodies[nd->body_index];
        if (b->mass <= 0.0)
            return;
        double bx = b->x - qx, by = b->y - qy, bz = b->z - qz;
        if (bx * bx + by * by + bz * bz <= r2)
            buf[(*count)++] = nd->body_index;
        return;
    }
    for (int c = 0; c < 8; ++c)
        if (nd->children[c] >= 0 && *count < cap)
            cref_nb_visit(pool, nd->children[c], bodies, qx, qy, qz, r2, buf, count, cap);
}

int cref_find_neighbors(const OctreeNode *pool, int root, const Body *bodies, int n, double qx,
                        double qy, double qz, double radius, int *buf, int cap)
{
    (void)n;
    int count = 0;
    cref_nb_visit(pool, root, bodies, qx, qy, qz, radius * radius, buf, &count, cap);
    return count;
}

/* ------------------------------------------------------------------ SPH
 * Smoothed-particle hydrodynamics of the gas bodies (src/sph.c:10-231): the M4
 * cubic-spline kernel (:12-41), density with the adaptive smoothing length
 * (:45-95), pressure gradient plus Monaghan viscosity (:99-196) and implicit
 * cooling (:200-231).  Every sum runs over the body's neighbour list (the
 * sphere query above, radius h, at most 128 bodies, in visiting order) with
 * the reference's expression order, so results are bit-identical.  The loops
 * are sequential: no body's result depends on another's within one pass. */
#define CREF_SPH_GAMMA (5.0 / 3.0)     /* src/sph.h:7 */
#define CREF_SPH_CAP 128               /* src/sph.h:9 */
#define CREF_PI 3.14159265358979323846

static double cref_spline(double r, double h)          /* src/sph.c:12-25 */
{
    if (h < 1e-10)
        return 0.0;
    double q = r / h;
    if (q > 1.0)
        return 0.0;
    double norm = 8.0 / (CREF_PI * h * h * h);
    if (q > 0.5) {
        double t = 1.0 - q;
        return norm * 2.0 * t * t * t;
    }
    return norm * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
}

static double cref_spline_dr(double r, double h)       /* src/sph.c:27-41 */
{
    if (h < 1e-10)
        return 0.0;
    double q = r / h;
    if (q > 1.0 || r < 1e-15)
        return 0.0;
    double norm = 8.0 / (CREF_PI * h * h * h);
    if (q > 0.5) {
        double t = 1.0 - q;
        return norm * (-6.0 * t * t) / h;
    }
    return norm * (-12.0 * q + 18.0 * q * q) / h;
}

static inline int cref_gas(const Body *b) { return b->type == BODY_GAS && !(b->mass <= 0.0); }

/* Both passes run their bodies under OpenMP as src/sph.c:50,106 does; each body
 * writes only its own fields and sums its neighbours in list order, so the
 * results do not depend on the thread count. */
void cref_sph_density(Body *bodies, int n, const OctreeNode *pool)
{
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i) {
        int list[CREF_SPH_CAP];
        Body *bi = &bodies[i];
        if (!cref_gas(bi))
            continue;
        const double h = bi->smoothing_h;
        const int k = cref_find_neighbors(pool, 0, bodies, n, bi->x, bi->y, bi->z, h, list,
                                          CREF_SPH_CAP);
        if (k < 20 && h < 50.0)                        /* :71-75 */
            bi->smoothing_h = h * 1.2;
        else if (k > 48 && h > 0.1)
            bi->smoothing_h = h * 0.8;
        double rho = 0.0;
        for (int a = 0; a < k; ++a) {
            const Body *bj = &bodies[list[a]];
            if (bj->mass <= 0.0)
                continue;
            double dx = bi->x - bj->x, dy = bi->y - bj->y, dz = bi->z - bj->z;
            rho += bj->mass * cref_spline(sqrt(dx * dx + dy * dy + dz * dz), h);
        }
        if (rho < 1e-15 || isnan(rho))
            rho = 1e-15;
        bi->density = rho;
        double u = bi->internal_energy;
        if (u < 1e-15)
            u = 1e-15;
        bi->pressure = (CREF_SPH_GAMMA - 1.0) * rho * u;
    }
}

void cref_sph_forces(Body *bodies, int n, const OctreeNode *pool)
{
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i) {
        int list[CREF_SPH_CAP];
        Body *bi = &bodies[i];
        if (!cref_gas(bi))
            continue;
        const double hi = bi->smoothing_h, rhoi = bi->density, Pi = bi->pressure;
        const double ci = sqrt(CREF_SPH_GAMMA * Pi / (rhoi + 1e-15));
        const int k = cref_find_neighbors(pool, 0, bodies, n, bi->x, bi->y, bi->z, hi, list,
                                          CREF_SPH_CAP);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int a = 0; a < k; ++a) {
            const int j = list[a];
            const Body *bj = &bodies[j];
            if (j == i || bj->mass <= 0.0 || bj->type != BODY_GAS)
                continue;
            double dx = bi->x - bj->x, dy = bi->y - bj->y, dz = bi->z - bj->z;
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r < 1e-15)
                continue;
            double hbar = 0.5 * (hi + bj->smoothing_h);
            double rhoj = bj->density, Pj = bj->pressure;
            double vr = (bi->vx - bj->vx) * dx + (bi->vy - bj->vy) * dy + (bi->vz - bj->vz) * dz;
            double visc = 0.0;
            if (vr < 0.0) {                            /* approaching: :152-161 */
                double mu = hbar * vr / (r * r + 0.01 * hbar * hbar);
                double cj = sqrt(CREF_SPH_GAMMA * Pj / (rhoj + 1e-15));
                double cbar = 0.5 * (ci + cj), rhobar = 0.5 * (rhoi + rhoj);
                visc = (-1.0 * cbar * mu + 2.0 * mu * mu) / (rhobar + 1e-15);
            }
            double g = cref_spline_dr(r, hbar);
            double gx = g * dx / r, gy = g * dy / r, gz = g * dz / r;
            double f = bj->mass * (Pi / (rhoi * rhoi + 1e-15) + Pj / (rhoj * rhoj + 1e-15) + visc);
            fx -= f * gx;
            fy -= f * gy;
            fz -= f * gz;
        }
        if (isnan(fx))
            fx = 0.0;
        if (isnan(fy))
            fy = 0.0;
        if (isnan(fz))
            fz = 0.0;
        bi->ax += fx * 0.1;                            /* :186-190 */
        bi->ay += fy * 0.1;
        bi->az += fz * 0.1;
    }
}

void cref_sph_cooling(Body *bodies, int n, double dt)
{
    for (int i = 0; i < n; ++i) {
        Body *b = &bodies[i];
        if (!cref_gas(b))
            continue;
        double u = b->internal_energy, rho = b->density;
        if (u < 1e-15 || rho < 1e-15)
            continue;
        double unew = u / (1.0 + dt * 0.05 * rho * 0.5 / sqrt(u));   /* :219-220 */
        if (unew < 0.01)
            unew = 0.01;
        b->internal_energy = unew;
        b->pressure = (CREF_SPH_GAMMA - 1.0) * rho * unew;
    }
}

/* ------------------------------------------------------------ quasar
 * One quasar physics step (src/quasar.c:477-510) and the compaction of dead
 * bodies (:512-523), with the random stream and the jets' knot phase as explicit
 * state instead of the reference's file-scope statics (quasar.c:11, :282).  The
 * sub-steps, their loops over bodies in index order and every expression keep
 * the reference's order, so results are bit-identical for the same state:
 *   accretion onto each black hole in index order (:113-198), whose mass,
 *   velocity and spin change inside the loop (:58-97); accretion rate and
 *   luminosity (:486-499); radiation feedback (:200-230); jet decay with lobes
 *   (:393-432); recycling of distant jets (:434-475); jet spawning (:232-391). */
typedef struct {               /* src/quasar.h:6-22 */
    double accretion_radius, swallow_radius, isco_radius, viscosity_alpha, jet_speed, jet_mass,
        jet_lifetime, feedback_strength, eta_eff, eddington_k, accretion_smoothing;
    int jet_cap, jet_count, max_bodies, jet_ring_count;
} CrefQuasarConfig;

static double crq_uniform(uint64_t *st)                 /* xorshift64, :14-27 */
{
    uint64_t x = *st;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *st = x;
    return (x >> 11) * (1.0 / 9007199254740992.0);
}

static double crq_gaussian(uint64_t *st)                /* Box-Muller, :29-36 */
{
    double u1 = crq_uniform(st);
    double u2 = crq_uniform(st);
    if (u1 < 1e-15)
        u1 = 1e-15;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * CREF_PI * u2);
}

static void crq_kill(Body *b)
{
    b->mass = 0.0;
    b->vx = b->vy = b->vz = 0.0;
    b->ax = b->ay = b->az = 0.0;
}

/* The black hole takes the victim's mass and momentum; its spin leans toward the
 * victim's orbital angular momentum (computed with the new velocity). */
static void crq_swallow(Body *bh, Body *v)              /* :58-97 */
{
    double M = bh->mass + v->mass;
    bh->vx = (bh->mass * bh->vx + v->mass * v->vx) / M;
    bh->vy = (bh->mass * bh->vy + v->mass * v->vy) / M;
    bh->vz = (bh->mass * bh->vz + v->mass * v->vz) / M;
    double rx = v->x - bh->x, ry = v->y - bh->y, rz = v->z - bh->z;
    double ux = v->vx - bh->vx, uy = v->vy - bh->vy, uz = v->vz - bh->vz;
    double lx = ry * uz - rz * uy, ly = rz * ux - rx * uz, lz = rx * uy - ry * ux;
    double l = sqrt(lx * lx + ly * ly + lz * lz);
    if (l > 1e-15) {
        double w = v->mass / M * 0.01;
        bh->spin_x += w * lx / l;
        bh->spin_y += w * ly / l;
        bh->spin_z += w * lz / l;
        double sm = sqrt(bh->spin_x * bh->spin_x + bh->spin_y * bh->spin_y +
                         bh->spin_z * bh->spin_z);
        if (sm > 1e-15) {
            bh->spin_x /= sm;
            bh->spin_y /= sm;
            bh->spin_z /= sm;
        }
    }
    bh->mass = M;
    crq_kill(v);
}

static void crq_drag(const Body *bh, Body *g, double alpha)   /* :99-121 */
{
    double rx = g->x - bh->x, ry = g->y - bh->y, rz = g->z - bh->z;
    double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1e-15)
        return;
    double ux = g->vx - bh->vx, uy = g->vy - bh->vy, uz = g->vz - bh->vz;
    double vr = (ux * rx + uy * ry + uz * rz) / r;
    g->ax -= alpha * (vr * rx / r);
    g->ay -= alpha * (vr * ry / r);
    g->az -= alpha * (vr * rz / r);
}

static double crq_accretion(Body *b, int n, const CrefQuasarConfig *q)   /* :123-198 */
{
    double swallowed = 0.0;
    const double racc2 = q->accretion_radius * q->accretion_radius;
    const double rsw2 = q->swallow_radius * q->swallow_radius;
    for (int s = 0; s < n; ++s) {
        if (b[s].type != BODY_SMBH || b[s].mass <= 0.0)
            continue;
        Body *bh = &b[s];
        for (int i = 0; i < n; ++i) {
            Body *v = &b[i];
            if (i == s || v->mass <= 0.0 || v->type == BODY_SMBH)
                continue;
            double dx = v->x - bh->x, dy = v->y - bh->y, dz = v->z - bh->z;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > racc2)
                continue;
            if (d2 < rsw2) {
                swallowed += v->mass;
                crq_swallow(bh, v);
                continue;
            }
            double r = sqrt(d2);
            double ux = v->vx - bh->vx, uy = v->vy - bh->vy, uz = v->vz - bh->vz;
            double lx = dy * uz - dz * uy, ly = dz * ux - dx * uz, lz = dx * uy - dy * ux;
            double rc = (lx * lx + ly * ly + lz * lz) / (bh->mass * r);
            if (rc > q->isco_radius && v->type != BODY_GAS) {
                v->type = BODY_GAS;
                if (v->smoothing_h < 1e-10)
                    v->smoothing_h = q->accretion_radius * 0.3;
                if (v->internal_energy < 1e-10)
                    v->internal_energy =
                        0.001 * (v->vx * v->vx + v->vy * v->vy + v->vz * v->vz) + 0.01;
            }
            if (v->type == BODY_GAS) {
                crq_drag(bh, v, q->viscosity_alpha);
                if (r < q->isco_radius) {
                    swallowed += v->mass;
                    crq_swallow(bh, v);
                }
            }
        }
    }
    return swallowed;
}

static void crq_feedback(Body *b, int n, const CrefQuasarConfig *q)     /* :200-230 */
{
    const double rfb2 = (q->accretion_radius * 3.0) * (q->accretion_radius * 3.0);
    for (int s = 0; s < n; ++s) {
        if (b[s].type != BODY_SMBH || b[s].luminosity <= 0.0)
            continue;
        const Body *bh = &b[s];
        for (int i = 0; i < n; ++i) {
            if (i == s || b[i].mass <= 0.0 || b[i].type == BODY_SMBH)
                continue;
            double dx = b[i].x - bh->x, dy = b[i].y - bh->y, dz = b[i].z - bh->z;
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > rfb2 || d2 < 1e-10)
                continue;
            double r = sqrt(d2);
            double f = q->feedback_strength * bh->luminosity / (4.0 * CREF_PI * d2);
            b[i].ax += f * dx / r;
            b[i].ay += f * dy / r;
            b[i].az += f * dz / r;
        }
    }
}

static void crq_decay(Body *b, int *n, int n_alloc, CrefQuasarConfig *q, double dt,
                      uint64_t *rng)                                      /* :393-432 */
{
    const int n0 = *n;
    for (int i = 0; i < n0; ++i) {
        if (b[i].type != BODY_JET || b[i].mass <= 0.0)
            continue;
        b[i].lifetime -= dt;
        if (b[i].lifetime > 0.0)
            continue;
        if (*n < n_alloc && b[i].mass > 0.2) {
            Body *lobe = &b[(*n)++];
            memset(lobe, 0, sizeof(*lobe));
            lobe->type = BODY_LOBE;
            lobe->mass = b[i].mass * 0.5;
            lobe->x = b[i].x;
            lobe->y = b[i].y;
            lobe->z = b[i].z;
            lobe->vx = b[i].vx * 0.15 + (crq_uniform(rng) - 0.5) * 0.5;
            lobe->vy = b[i].vy * 0.15 + (crq_uniform(rng) - 0.5) * 0.5;
            lobe->vz = b[i].vz * 0.15 + (crq_uniform(rng) - 0.5) * 0.5;
            lobe->lifetime = q->jet_lifetime * 3.0;
        }
        crq_kill(&b[i]);
        q->jet_count--;
    }
    for (int i = 0; i < *n; ++i) {
        if (b[i].type != BODY_LOBE || b[i].mass <= 0.0)
            continue;
        b[i].lifetime -= dt;
        if (b[i].lifetime <= 0.0)
            crq_kill(&b[i]);
    }
}

static void crq_recycle(Body *b, int n, CrefQuasarConfig *q, uint64_t *rng)   /* :434-475 */
{
    int s = 0;
    while (s < n && !(b[s].type == BODY_SMBH && b[s].mass > 0.0))
        ++s;
    if (s == n)
        return;
    const Body *bh = &b[s];
    const double rr = q->accretion_radius * 10.0, rr2 = rr * rr;
    for (int i = 0; i < n; ++i) {
        Body *j = &b[i];
        if (j->type != BODY_JET || j->mass <= 0.0)
            continue;
        double dx = j->x - bh->x, dy = j->y - bh->y, dz = j->z - bh->z;
        if (dx * dx + dy * dy + dz * dz < rr2)
            continue;
        double angle = 2.0 * CREF_PI * crq_uniform(rng);
        double cphi = 2.0 * crq_uniform(rng) - 1.0;
        double sphi = sqrt(1.0 - cphi * cphi);
        double rp = q->accretion_radius;
        j->x = bh->x + rp * sphi * cos(angle);
        j->y = bh->y + rp * sphi * sin(angle);
        j->z = bh->z + rp * cphi;
        double vin = 0.5 * sqrt(bh->mass / rp);
        double nx = (j->x - bh->x) / rp, ny = (j->y - bh->y) / rp, nz = (j->z - bh->z) / rp;
        j->vx = bh->vx - vin * nx;
        j->vy = bh->vy - vin * ny;
        j->vz = bh->vz - vin * nz;
        j->ax = j->ay = j->az = 0.0;
        j->type = BODY_GAS;
        j->lifetime = 0.0;
        j->smoothing_h = q->accretion_radius * 0.3;
        j->internal_energy = 0.001 * (j->vx * j->vx + j->vy * j->vy + j->vz * j->vz) + 0.01;
        j->density = 0.0;
        j->pressure = 0.0;
        q->jet_count--;
    }
}

static void crq_spawn(Body *b, int *n, int n_alloc, CrefQuasarConfig *q, uint64_t *rng,
                      int *phase)                                         /* :232-391 */
{
    const int k = ++(*phase);
    double smod = 1.0 + 0.6 * sin(k * 0.10);
    if (smod < 0.5)
        smod = 0.5;
    const double cone = 0.05 + 0.13 * (0.5 + 0.5 * sin(k * 0.057));
    const double hph = k * 0.04;
    for (int s = 0; s < *n; ++s) {
        if (b[s].type != BODY_SMBH || b[s].luminosity <= 0.0)
            continue;
        const Body *bh = &b[s];
        const double energy = 0.5 * q->jet_mass * q->jet_speed * q->jet_speed;
        if (energy < 1e-15)
            continue;
        int pulses = (int)(bh->luminosity / energy);
        if (pulses < 2)
            pulses = 2;
        if (pulses > 8)
            pulses = 8;
        /* two unit vectors across the spin axis */
        double px, py, pz;
        if (fabs(bh->spin_z) < 0.9) {
            px = -bh->spin_y; py = bh->spin_x; pz = 0.0;
        } else {
            px = 0.0; py = -bh->spin_z; pz = bh->spin_y;
        }
        double pm = sqrt(px * px + py * py + pz * pz);
        if (pm > 1e-15) {
            px /= pm; py /= pm; pz /= pm;
        }
        double qx = bh->spin_y * pz - bh->spin_z * py;
        double qy = bh->spin_z * px - bh->spin_x * pz;
        double qz = bh->spin_x * py - bh->spin_y * px;
        double qm = sqrt(qx * qx + qy * qy + qz * qz);
        if (qm > 1e-15) {
            qx /= qm; qy /= qm; qz /= qm;
        }
        const double ha = q->swallow_radius * 0.35;
        const double hx = ha * (cos(hph) * px + sin(hph) * qx);
        const double hy = ha * (cos(hph) * py + sin(hph) * qy);
        const double hz = ha * (cos(hph) * pz + sin(hph) * qz);
        const double speed = q->jet_speed * smod;
        const int total = pulses + ((smod > 1.35) ? 4 : 0);
        for (int j = 0; j < total && q->jet_count < q->jet_cap; ++j) {
            if (*n >= n_alloc)
                break;
            const int burst = j >= pulses;
            const double sg = (j % 2 == 0) ? 1.0 : -1.0;
            const double cn = burst ? 0.012 : cone;
            const double sp = burst ? speed * 1.28 : speed;
            const double ms = burst ? q->jet_mass * 1.9 : q->jet_mass;
            const double life = burst ? q->jet_lifetime * 1.4 : q->jet_lifetime;
            const double fwd = burst ? q->swallow_radius * 0.6 : 0.0;
            const double ang = crq_uniform(rng) * 2.0 * CREF_PI;
            const double off = crq_gaussian(rng) * cn;
            const int ring = burst ? 1 : q->jet_ring_count;
            for (int r = 0; r < ring; ++r) {
                if (*n >= n_alloc || q->jet_count >= q->jet_cap)
                    break;
                Body *e = &b[*n];
                memset(e, 0, sizeof(*e));
                e->type = BODY_JET;
                e->lifetime = life;
                double ra = (ring > 1) ? 2.0 * CREF_PI * r / ring + ang : ang;
                double rr = (ring > 1) ? q->swallow_radius * 0.15 : 0.0;
                double ox = rr * (cos(ra) * px + sin(ra) * qx);
                double oy = rr * (cos(ra) * py + sin(ra) * qy);
                double oz = rr * (cos(ra) * pz + sin(ra) * qz);
                e->mass = (ring > 1) ? ms * 0.6 : ms;
                e->x = bh->x + sg * bh->spin_x * (q->swallow_radius + fwd) + off * (px * cos(ang)) +
                       sg * hx + ox;
                e->y = bh->y + sg * bh->spin_y * (q->swallow_radius + fwd) + off * (py * cos(ang)) +
                       sg * hy + oy;
                e->z = bh->z + sg * bh->spin_z * (q->swallow_radius + fwd) + off * (pz * sin(ang)) +
                       sg * hz + oz;
                double ev = sp * 0.04 * (1.0 + e->lifetime * 0.1);
                e->vx = bh->vx + sg * bh->spin_x * sp + ox * ev / (rr + 1e-15);
                e->vy = bh->vy + sg * bh->spin_y * sp + oy * ev / (rr + 1e-15);
                e->vz = bh->vz + sg * bh->spin_z * sp + oz * ev / (rr + 1e-15);
                (*n)++;
                q->jet_count++;
            }
        }
    }
}

void cref_quasar_step(Body *b, int *n, int n_alloc, CrefQuasarConfig *q, double dt,
                      uint64_t *rng, int *phase)
{
    const double swallowed = crq_accretion(b, *n, q);
    for (int i = 0; i < *n; ++i) {                                       /* :486-499 */
        if (b[i].type != BODY_SMBH)
            continue;
        const double mdot = swallowed / dt;
        b[i].accretion_rate = q->accretion_smoothing * b[i].accretion_rate +
                              (1.0 - q->accretion_smoothing) * mdot;
        const double ledd = q->eddington_k * b[i].mass;
        b[i].luminosity = q->eta_eff * b[i].accretion_rate;
        if (b[i].luminosity > ledd)
            b[i].luminosity = ledd;
    }
    crq_feedback(b, *n, q);
    crq_decay(b, n, n_alloc, q, dt, rng);
    crq_recycle(b, *n, q, rng);
    crq_spawn(b, n, n_alloc, q, rng, phase);
}

void cref_decay_jets(Body *b, int *n, int n_alloc, CrefQuasarConfig *q, double dt,
                     uint64_t *rng)
{
    crq_decay(b, n, n_alloc, q, dt, rng);
}

int cref_quasar_compact(Body *b, int n)                                  /* :512-523 */
{
    int w = 0;
    for (int r = 0; r < n; ++r)
        if (b[r].mass > 0.0)
            b[w++] = b[r];
    return w;
}

/* ------------------------------------------------------------ snapshots
 * The host consumers' float snapshots of the live bodies in index order:
 * the Unreal plugin's FCosmosimBodyGPU, 16 floats, at most max_out records
 * (ue/CosmosimPlugin/.../CosmosimSubsystem.cpp:12-43), and the renderer's
 * upload record, 10 floats (src/renderer.c:348-364).  Returns the count. */
int cref_snapshot(const Body *bodies, int n, int fmt, float *out, int max_out)
{
    int k = 0;
    for (int i = 0; i < n && k < max_out; ++i) {
        const Body *b = &bodies[i];
        if (b->mass <= 0.0)
            continue;
        float *o = out + (size_t)k * (fmt == 0 ? 16 : 10);
        o[0] = (float)b->x; o[1] = (float)b->y; o[2] = (float)b->z;
        o[3] = (float)b->mass;
        o[4] = (float)b->vx; o[5] = (float)b->vy; o[6] = (float)b->vz;
        o[7] = (float)b->type;
        o[8] = (float)b->internal_energy;
        o[9] = (float)b->density;
        if (fmt == 0) {
            o[10] = (float)b->lifetime;
            o[11] = (float)b->luminosity;
            o[12] = (float)b->spin_x; o[13] = (float)b->spin_y; o[14] = (float)b->spin_z;
            o[15] = 0.0f;
        }
        ++k;
    }
    return k;
}

/* Reference-layout pools differ between builders only in node numbering when
 * the trees are equal.  cref_canonical renumbers a pool in depth-first
 * pre-order (children in slot order) into out[], so two pools can be compared
 * field by field.  Returns the number of nodes reached from root. */
int cref_canonical(const OctreeNode *pool, int root, OctreeNode *out, int *map, int pool_size)
{
    for (int i = 0; i < pool_size; ++i)
        map[i] = -1;
    int *stack = (int *)malloc(sizeof(int) * (size_t)(pool_size + 8));
    int top = 0, next = 0;
    stack[top++] = root;
    while (top > 0) {
        int v = stack[--top];
        map[v] = next++;
        for (int c = 7; c >= 0; --c)
            if (pool[v].children[c] >= 0)
                stack[top++] = pool[v].children[c];
    }
    for (int v = 0; v < pool_size; ++v) {
        if (map[v] < 0)
            continue;
        OctreeNode nd = pool[v];
        for (int c = 0; c < 8; ++c)
            if (nd.children[c] >= 0)
                nd.children[c] = map[nd.children[c]];
        out[map[v]] = nd;
    }
    free(stack);
    return next;
}
