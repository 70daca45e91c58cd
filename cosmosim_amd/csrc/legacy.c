/*
 * legacy.c - the reference's L1 entry points (src/octree.h:16-36,
 * src/integrator.h:7-17) on top of the HIP core.  Host code stays C: it moves
 * Body[] and pools across the C-ABI in include/csg.h and lets the device do the
 * build, the walk and the kicks.  The caller's Body[] stays the source of truth
 * between calls, as src/quasar.c and src/sph.c expect (they change n, masses,
 * positions and accelerations between calls).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_common.h"
#include "integrator.h"
#include "octree.h"

/* ------------------------------------------------------------ errors, ctx */
static _Thread_local int tl_error = 0;
static _Thread_local csg_ctx *tl_ctx = NULL;

void cs_set_error(int code, const char *where)
{
    tl_error = code;
    if (code != COSMOSIM_OK)
        fprintf(stderr, "cosmosim: %s failed: %s\n", where, cosmosim_ext_error_string(code));
}

int cosmosim_ext_last_error(void) { return tl_error; }
void cosmosim_ext_clear_error(void) { tl_error = COSMOSIM_OK; }

int cs_default_device(void)
{
    const char *s = getenv("COSMOSIM_DEVICE");
    return s ? atoi(s) : 0;
}

csg_ctx *cs_thread_ctx(void)
{
    if (!tl_ctx) {
        int st = csg_create(cs_default_device(), 0, &tl_ctx);
        if (st != COSMOSIM_OK) {
            cs_set_error(st, "device context");
            tl_ctx = NULL;
        }
    }
    return tl_ctx;
}

/* ------------------------------------------------------------ body transfer */
int cs_download_kinematics(csg_ctx *ctx, Body *bodies, int n, int pos, int vel, int acc)
{
    if (n <= 0)
        return COSMOSIM_OK;
    double *buf = (double *)malloc(sizeof(double) * 3 * (size_t)n * 3);
    if (!buf)
        return COSMOSIM_E_ALLOC;
    double *p = buf, *v = buf + 3 * (size_t)n, *a = buf + 6 * (size_t)n;
    int st = csg_get_state(ctx, pos ? p : NULL, vel ? v : NULL, acc ? a : NULL, NULL);
    if (st == COSMOSIM_OK) {
        for (int i = 0; i < n; ++i) {
            Body *b = &bodies[i];
            if (pos) { b->x = p[3 * i]; b->y = p[3 * i + 1]; b->z = p[3 * i + 2]; }
            if (vel) { b->vx = v[3 * i]; b->vy = v[3 * i + 1]; b->vz = v[3 * i + 2]; }
            if (acc) { b->ax = a[3 * i]; b->ay = a[3 * i + 1]; b->az = a[3 * i + 2]; }
        }
    }
    free(buf);
    return st;
}

/* FNV-1a over the bytes of x, y, z and mass. */
uint64_t cs_body_hash(const Body *bodies, int n)
{
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < n; ++i) {
        const double f[4] = {bodies[i].x, bodies[i].y, bodies[i].z, bodies[i].mass};
        const unsigned char *c = (const unsigned char *)f;
        for (int k = 0; k < 32; ++k) {
            h ^= c[k];
            h *= 1099511628211ull;
        }
    }
    return h ^ (uint64_t)(unsigned)n;
}

/* The tree last written by octree_build / integrator_* on this thread. */
static _Thread_local struct {
    const OctreeNode *pool;
    int pool_size;
    int n;
    uint64_t hash;
    OctreeNode root;
    int valid;
} tl_tree;

void cs_note_tree(const OctreeNode *pool, int pool_size, const Body *bodies, int n)
{
    tl_tree.pool = pool;
    tl_tree.pool_size = pool_size;
    tl_tree.n = n;
    tl_tree.hash = cs_body_hash(bodies, n);
    memcpy(&tl_tree.root, &pool[0], sizeof(OctreeNode));
    tl_tree.valid = 1;
}

void cs_forget_tree(void) { tl_tree.valid = 0; }

int cs_tree_matches(const OctreeNode *pool, int root, const Body *bodies, int n)
{
    return tl_tree.valid && root == 0 && pool == tl_tree.pool && n == tl_tree.n &&
           memcmp(&pool[0], &tl_tree.root, sizeof(OctreeNode)) == 0 &&
           cs_body_hash(bodies, n) == tl_tree.hash;
}

int cs_pool_extent(const OctreeNode *pool, int root)
{
    /* iterative DFS with a growable stack */
    int cap = 1024, top = 0, hi = root;
    int *st = (int *)malloc(sizeof(int) * (size_t)cap);
    if (!st)
        return -1;
    st[top++] = root;
    while (top > 0) {
        int v = st[--top];
        if (v > hi)
            hi = v;
        for (int c = 0; c < 8; ++c) {
            int ch = pool[v].children[c];
            if (ch < 0)
                continue;
            if (top == cap) {
                int *g = (int *)realloc(st, sizeof(int) * (size_t)cap * 2);
                if (!g) { free(st); return -1; }
                st = g;
                cap *= 2;
            }
            st[top++] = ch;
        }
    }
    free(st);
    return hi + 1;
}

static int pool_capacity(int n) { return 8 * (n > 1 ? n : 1); }

/* ------------------------------------------------------------ substeps */
int cs_legacy_init(csg_ctx *ctx, Body *bodies, int n, double G, double softening, double theta,
                   OctreeNode *pool, int pool_cap)
{
    int st, used = 0;
    if ((st = csg_set_bodies(ctx, bodies, n)))
        return st;
    if ((st = csg_build(ctx)))
        return st;
    if ((st = csg_walk(ctx, COSMOSIM_PREC_FP64_STRICT, G, softening, theta)))
        return st;
    if (pool && (st = csg_export_pool(ctx, pool, pool_cap, &used)))
        return st;
    if ((st = cs_download_kinematics(ctx, bodies, n, 0, 0, 1)))
        return st;
    if (pool)
        cs_note_tree(pool, used, bodies, n);
    return COSMOSIM_OK;
}

int cs_legacy_step(csg_ctx *ctx, Body *bodies, int n, double dt, double G, double softening,
                   double theta, OctreeNode *pool, int pool_cap, int sph_enabled)
{
    int st, used = 0;
    if ((st = csg_set_bodies(ctx, bodies, n)))
        return st;
    if ((st = csg_kick_drift(ctx, dt)))
        return st;
    if ((st = csg_build(ctx)))
        return st;
    if ((st = csg_walk(ctx, COSMOSIM_PREC_FP64_STRICT, G, softening, theta)))
        return st;
    if (pool && (st = csg_export_pool(ctx, pool, pool_cap, &used)))
        return st;
    if (sph_enabled) {
        /* SPH on the device between the walk and the second half kick
         * (src/integrator.c:39-42), cooling after it (:54-56); the download
         * returns the gas fields with the kinematics */
        if ((st = csg_sph_density(ctx)) || (st = csg_sph_forces(ctx)))
            return st;
        if ((st = csg_kick(ctx, dt)) || (st = csg_sph_cooling(ctx, dt)))
            return st;
        if ((st = csg_get_bodies(ctx, bodies, n)))
            return st;
        if (pool)
            cs_note_tree(pool, used, bodies, n);
        return COSMOSIM_OK;
    }
    if ((st = csg_kick(ctx, dt)))
        return st;
    if ((st = cs_download_kinematics(ctx, bodies, n, 1, 1, 1)))
        return st;
    if (pool)
        cs_note_tree(pool, used, bodies, n);
    return COSMOSIM_OK;
}

/* ------------------------------------------------------------ src/octree.h */
void octree_build(OctreeNode *pool, int *pool_size, const Body *bodies, int n)
{
    if (pool_size)
        *pool_size = 0;
    cs_forget_tree();
    if (!pool || !pool_size || n < 0 || (n > 0 && !bodies)) {
        cs_set_error(COSMOSIM_E_ARG, "octree_build");
        return;
    }
    csg_ctx *ctx = cs_thread_ctx();
    if (!ctx)
        return;
    int st, used = 0;
    if ((st = csg_set_bodies(ctx, bodies, n)) || (st = csg_build(ctx)) ||
        (st = csg_export_pool(ctx, pool, pool_capacity(n), &used))) {
        cs_set_error(st, "octree_build");
        return;
    }
    *pool_size = used;
    cs_note_tree(pool, used, bodies, n);
    cs_set_error(COSMOSIM_OK, "octree_build");
}

void octree_compute_forces(const OctreeNode *pool, int root, Body *bodies, int n, double G,
                           double softening, double theta)
{
    if (!pool || n < 0 || (n > 0 && !bodies)) {
        cs_set_error(COSMOSIM_E_ARG, "octree_compute_forces");
        return;
    }
    csg_ctx *ctx = cs_thread_ctx();
    if (!ctx)
        return;
    int st;
    if (cs_tree_matches(pool, root, bodies, n)) {
        /* the device already holds this tree and these bodies */
        st = csg_walk(ctx, COSMOSIM_PREC_FP64_STRICT, G, softening, theta);
    } else {
        int extent = cs_pool_extent(pool, root);
        st = extent < 0 ? COSMOSIM_E_ALLOC : csg_set_bodies(ctx, bodies, n);
        if (st == COSMOSIM_OK)
            st = csg_walk_pool(ctx, pool, extent, root, G, softening, theta);
    }
    if (st == COSMOSIM_OK)
        st = cs_download_kinematics(ctx, bodies, n, 0, 0, 1);
    cs_set_error(st, "octree_compute_forces");
}

/* Sphere query for the SPH consumer (src/octree.c:236-318).  It runs on the
 * host because its callers (src/sph.c:57,120) issue one small query per
 * particle from host OpenMP loops over a host pool; the gravity hot path does
 * not use it, and batched queries run on the device
 * (cosmosim_ext_find_neighbors).
 *
 * Depth-first over an explicit stack: children are pushed in reverse slot
 * order, so they pop in slot order, which is the order the reported bodies
 * must follow (the cap cuts the list).  A node whose cell misses the sphere is
 * dropped when popped; a leaf reports its body when it is live and inside.
 * Once the list is full nothing more can be reported, so the walk stops. */
#define NB_STACK (8 * 65 + 8) /* 8 children per level, depth <= 64 */

static int cell_misses_sphere(const OctreeNode *nd, double qx, double qy, double qz, double r2)
{
    const double gx = fmax(0.0, fabs(qx - nd->center_x) - nd->half_size);
    const double gy = fmax(0.0, fabs(qy - nd->center_y) - nd->half_size);
    const double gz = fmax(0.0, fabs(qz - nd->center_z) - nd->half_size);
    return gx * gx + gy * gy + gz * gz > r2;
}

int octree_find_neighbors(const OctreeNode *pool, int root, const Body *bodies, int n, double cx,
                          double cy, double cz, double radius, int *neighbor_buf,
                          int buf_capacity)
{
    (void)n;
    if (!pool || !bodies || !neighbor_buf)
        return 0;
    const double r2 = radius * radius;
    int stack[NB_STACK];
    int top = 0, found = 0;
    stack[top++] = root;
    while (top > 0 && found < buf_capacity) {
        const OctreeNode *nd = &pool[stack[--top]];
        if (cell_misses_sphere(nd, cx, cy, cz, r2))
            continue;
        const int bi = nd->body_index;
        if (bi >= 0) {
            const Body *b = &bodies[bi];
            if (b->mass > 0.0) {
                const double ox = b->x - cx, oy = b->y - cy, oz = b->z - cz;
                if (ox * ox + oy * oy + oz * oz <= r2)
                    neighbor_buf[found++] = bi;
            }
            continue;
        }
        for (int slot = 7; slot >= 0; --slot) {
            const int child = nd->children[slot];
            if (child >= 0 && top < NB_STACK)
                stack[top++] = child;
        }
    }
    return found;
}

/* ------------------------------------------------------------ src/integrator.h */
void integrator_init_accelerations(Body *bodies, int n, double G, double softening, double theta,
                                   OctreeNode *pool)
{
    cs_forget_tree();
    if (n < 0 || (n > 0 && !bodies)) {
        cs_set_error(COSMOSIM_E_ARG, "integrator_init_accelerations");
        return;
    }
    csg_ctx *ctx = cs_thread_ctx();
    if (!ctx)
        return;
    int st = cs_legacy_init(ctx, bodies, n, G, softening, theta, pool, pool_capacity(n));
    cs_set_error(st, "integrator_init_accelerations");
}

void integrator_step(Body *bodies, int n, double dt, double G, double softening, double theta,
                     OctreeNode *pool, int sph_enabled)
{
    cs_forget_tree();
    if (n < 0 || (n > 0 && !bodies)) {
        cs_set_error(COSMOSIM_E_ARG, "integrator_step");
        return;
    }
    csg_ctx *ctx = cs_thread_ctx();
    if (!ctx)
        return;
    int st = cs_legacy_step(ctx, bodies, n, dt, G, softening, theta, pool, pool_capacity(n),
                            sph_enabled);
    cs_set_error(st, "integrator_step");
}
