/*
 * csg.h - the thin C-ABI between the C host layer and the HIP gravity core
 * ("cosmosim gravity", gfx950).  No counterpart in the reference: it is the
 * seam below src/octree.c and src/integrator.c, whose work it carries out on
 * one GPU.  Plain pointers and sizes only; every call returns 0 or a negative
 * COSMOSIM_E_* code from cosmosim_ext.h.
 *
 * A csg_ctx owns one HIP stream and the device state of up to `capacity`
 * bodies: FP64 SoA arrays x,y,z,vx,vy,vz,ax,ay,az,m in the caller's body order,
 * the sort/tree scratch and the tree itself.  All calls on one context must
 * come from one thread at a time; each call makes the context's device current
 * on the calling thread.
 */
#ifndef CSG_H
#define CSG_H

#include "body.h"
#include "octree.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csg_ctx csg_ctx;

typedef struct {
    int n, n_live, n_nodes, max_depth, collision_runs, overflow;
    long long interactions;
    double root[4];                 /* centre x, y, z and half size of the root cube */
    double ms_build, ms_walk, ms_kdk; /* GPU time per phase since timing was (re)enabled */
    int timed_walks;                /* walk launches covered by ms_walk */
    int list_fallbacks;             /* interaction-list walk wavefronts that fell back to
                                       the union traversal since timing was (re)enabled */
    double ms_neighbors;            /* GPU time of neighbour-query traversals (timing on) */
    long long neighbor_queries;     /* queries covered by ms_neighbors */
    double ms_sph;                  /* GPU time of SPH density + force passes (timing on) */
} csg_stats;

int  csg_device_count(int *count);
int  csg_device_info(int device, char *name, int len, int *compute_units);
int  csg_device_memory(int device, unsigned long long *free_bytes,
                       unsigned long long *total_bytes);
int  csg_create(int device, int capacity, csg_ctx **out);
void csg_destroy(csg_ctx *ctx);
int  csg_reserve(csg_ctx *ctx, int capacity);
int  csg_size(const csg_ctx *ctx);

/* Host AoS <-> device.  set_bodies copies all 168-byte records (the fields the
 * core does not touch are kept on the device and written back unchanged by
 * get_bodies).  set_accelerations uploads only ax, ay, az. */
int  csg_set_bodies(csg_ctx *ctx, const Body *bodies, int n);
int  csg_get_bodies(csg_ctx *ctx, Body *bodies, int n);
int  csg_set_accelerations(csg_ctx *ctx, const Body *bodies, int n);
int  csg_get_state(csg_ctx *ctx, double *pos, double *vel, double *acc, double *mass);
int  csg_set_state(csg_ctx *ctx, const double *pos, const double *vel,
                   const double *acc, const double *mass);

/* KDK pieces (src/integrator.c:22-31 and :45-51). */
int  csg_kick_drift(csg_ctx *ctx, double dt);
int  csg_kick(csg_ctx *ctx, double dt);
/* Tree over the live bodies at the current positions (src/octree.c:114-169). */
int  csg_build(csg_ctx *ctx);
/* Force walk over the last tree (src/octree.c:171-234); precision is a
 * COSMOSIM_PREC_* value.  Dead bodies get a = 0. */
int  csg_walk(csg_ctx *ctx, int precision, double G, double softening, double theta);
/* k full substeps on the device: kick+drift, build, walk, kick (src/integrator.c:12-57
 * without the SPH hooks). */
int  csg_substeps(csg_ctx *ctx, int k, double dt, double G, double softening,
                  double theta, int precision);
/* Last tree in reference layout (root 0); fails with COSMOSIM_E_CAPACITY when
 * the tree has more than capacity nodes. */
int  csg_export_pool(csg_ctx *ctx, OctreeNode *pool, int capacity, int *pool_size);
/* Sphere queries over the current tree (batched GPU form of src/octree.c:236-318):
 * centres [k][3], radii [k]; out [k][cap] receives the first counts[q] <= cap
 * reported bodies of query q in the reference's visiting order. */
int  csg_find_neighbors(csg_ctx *ctx, const double *centres, const double *radius, int k,
                        int cap, int *out, int *counts);
/* SPH of the gas bodies on the device state and the current tree: the GPU forms
 * of sph_compute_density, sph_compute_forces and sph_apply_cooling
 * (src/sph.h:11-20), bit-identical to src/sph.c on the same bodies and tree.
 * Type, smoothing length, density, pressure and internal energy live in the
 * device body records; downloads (csg_get_bodies) return them. */
int  csg_sph_density(csg_ctx *ctx);
int  csg_sph_forces(csg_ctx *ctx);
int  csg_sph_cooling(csg_ctx *ctx, double dt);
/* k device-resident substeps of integrator_step with SPH enabled
 * (src/integrator.c:12-57). */
int  csg_substeps_sph(csg_ctx *ctx, int k, double dt, double G, double softening, double theta,
                      int precision);
/* Host snapshot formats written on the device from the device state: the live
 * bodies in index order as floats.  fmt 0: the Unreal plugin's FCosmosimBodyGPU
 * (16 floats, at most max_bodies records); fmt 1: the renderer's upload record
 * (10 floats).  *count = records copied to out. */
int  csg_snapshot(csg_ctx *ctx, int fmt, float *out, int max_bodies, int *count);
/* Quasar physics on the device state.  csg_quasar_config has the byte layout of
 * the reference's QuasarConfig (src/quasar.h:6-22). */
typedef struct {
    double accretion_radius, swallow_radius, isco_radius, viscosity_alpha, jet_speed, jet_mass,
        jet_lifetime, feedback_strength, eta_eff, eddington_k, accretion_smoothing;
    int jet_cap, jet_count, max_bodies, jet_ring_count;
} csg_quasar_config;
/* quasar_step (src/quasar.c:477-510) on the device bodies: accretion, feedback,
 * jet decay and recycling, jet spawning.  cfg->jet_count is updated; the body
 * count (csg_size) may grow up to the context's capacity.  The random stream and
 * the knot phase are the context's (csg_quasar_state). */
int  csg_quasar_step(csg_ctx *ctx, csg_quasar_config *cfg, double dt);
/* quasar_compact (src/quasar.c:512-523): live bodies keep their order. */
int  csg_quasar_compact(csg_ctx *ctx, int *n_out);
/* Get (set = 0) or set (set = 1) the quasar random state and knot phase; a new
 * context starts at 1 and 0, as the reference's statics. */
int  csg_quasar_state(csg_ctx *ctx, unsigned long long *rng, int *phase, int set);
/* Walk a caller-supplied reference-layout pool from `root` (exact reference
 * traversal, FP64 strict arithmetic). */
int  csg_walk_pool(csg_ctx *ctx, const OctreeNode *pool, int pool_size, int root,
                   double G, double softening, double theta);
/* Diagnostics. */
int  csg_direct_forces(csg_ctx *ctx, const int *idx, int k, double G, double softening,
                       double *acc);
int  csg_energy(csg_ctx *ctx, double G, double softening, double *kinetic,
                double *potential, double momentum[3]);
int  csg_interaction_counts(csg_ctx *ctx, int *counts, int n);
int  csg_get_stats(csg_ctx *ctx, csg_stats *out);
/* Enable (and reset) per-phase GPU timing: hipEvents on the context's stream,
 * recorded without host waits and collected by csg_get_stats. */
int  csg_set_timing(csg_ctx *ctx, int enabled);
int  csg_synchronize(csg_ctx *ctx);
/* Test hook: the nth build from now (1 = the next) runs with no node capacity,
 * so it fails as a tree that does not fit would; 0 cancels. */
int  csg_debug_fail_build(csg_ctx *ctx, int nth);
/* Test hook: the last tree's body orders, n_live entries each (any output may be
 * NULL): the sorted order (body indices by octree key), the wavefront order the
 * last build computed for the untaped walks (the same bodies in Peano-Hilbert order
 * of the cells; the sorted order when the build computed none: fewer than 16384
 * bodies, or an FP32 substep build with no more wavefronts than the GPU holds at
 * once; the environment variables COSMOSIM_HILBERT_MIN_BODIES and
 * COSMOSIM_HILBERT_MIN_WAVES at context creation change both thresholds) and the
 * sorted 48-bit keys (16 levels of octant digits, the first level in bits 47..45).
 * COSMOSIM_E_CAPACITY when n < n_live (*n_live is set first). */
int  csg_debug_orders(csg_ctx *ctx, int *sorted_idx, int *wave_idx, unsigned long long *keys,
                      int n, int *n_live);

/* Adjoint (no counterpart in the reference).
 * csg_walk_taped: the walk, also recording its tape (the visited nodes and the
 *   interacting bodies of each wavefront): frozen tree + opening decisions.
 * csg_force_vjp: xbar = J^T w for a = forces(x) at the last taped walk, w and
 *   xbar [n][3] in body order (centre-of-mass dependence included, masses fixed).
 * csg_forces_frozen: forces at positions pos (NULL: current) through that tree
 *   and tape, centres of mass from pos; pos becomes the state's positions.
 * csg_direct_vjp: the same VJP for the FP64 direct sum (the judge).
 * csg_record: k substeps from the current state keeping a checkpoint every
 *   `every` substeps (and, with keep_trees, every force evaluation's tree+tape).
 * csg_backward: dL/dx_0, dL/dv_0 from dL/dx_k, dL/dv_k (gv may be NULL) through
 *   the recorded run; mode 0 recomputes segments from checkpoints, mode 1
 *   reverses the leapfrog from the final state.  stats: [0] max |x| deviation of
 *   the reversed states from the checkpoints (mode 1), [1] adjoint device bytes.
 * csg_replay: frozen-tree forward of the recorded run from (pos0, vel0). */
int  csg_walk_taped(csg_ctx *ctx, int precision, double G, double softening, double theta);
int  csg_force_vjp(csg_ctx *ctx, const double *w, double *xbar);
int  csg_forces_frozen(csg_ctx *ctx, const double *pos, double *acc);
int  csg_direct_vjp(csg_ctx *ctx, const double *w, double *xbar, double G, double softening);
int  csg_record(csg_ctx *ctx, int k, double dt, double G, double softening, double theta,
                int precision, int every, int keep_trees);
int  csg_backward(csg_ctx *ctx, const double *gx, const double *gv, int mode, double *x0bar,
                  double *v0bar, double *stats);
int  csg_replay(csg_ctx *ctx, const double *pos0, const double *vel0, double *pos, double *vel);
int  csg_set_tape(csg_ctx *ctx, int enabled);

/* Pinned host memory for the handle-owned Body[] (zero-filled). */
void *csg_host_alloc(size_t bytes);
void  csg_host_free(void *p);

#ifdef __cplusplus
}
#endif

#endif
