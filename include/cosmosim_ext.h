/*
 * cosmosim_ext.h - extension entry points of libcosmosim (no counterpart in the
 * reference).  They expose what the eight legacy exports cannot: caller-chosen
 * G / softening / theta / precision, device-resident stepping without a host
 * round trip per substep, diagnostics (direct-sum judge, energy, momentum,
 * interaction counts, tree export) and a seeded initial-condition generator.
 *
 * Every function returns an int status: COSMOSIM_OK (0) or a negative
 * COSMOSIM_E_* code.  Handles from cosmosim_create() and cosmosim_ext_create()
 * are interchangeable: the eight legacy getters work on both.
 */
#ifndef COSMOSIM_EXT_H
#define COSMOSIM_EXT_H

#include "body.h"
#include "octree.h"
#include "cosmosim_api.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    COSMOSIM_OK = 0,
    COSMOSIM_E_ARG = -1,        /* invalid argument (NULL pointer, n < 0, bad enum) */
    COSMOSIM_E_NO_DEVICE = -2,  /* no usable HIP device */
    COSMOSIM_E_ALLOC = -3,      /* host or device allocation failed */
    COSMOSIM_E_HIP = -4,        /* a HIP runtime call or kernel launch failed */
    COSMOSIM_E_CAPACITY = -5,   /* tree needs more than 8*n nodes, or an output buffer is too small */
    COSMOSIM_E_STATE = -6,      /* call is not valid in the handle's current state */
    COSMOSIM_E_UNSUPPORTED = -7 /* feature not available in this build */
};

/* Arithmetic of the force walk. */
enum {
    COSMOSIM_PREC_FP64_STRICT = 0, /* reference arithmetic: IEEE sqrt and divide, no FMA contraction */
    COSMOSIM_PREC_FP64 = 1,        /* FP64 with refined hardware rsqrt and FMA */
    COSMOSIM_PREC_FP32 = 2,        /* FP32 on FP64-formed node offsets, compensated accumulation */
    COSMOSIM_PREC_FP32_LIST = 3    /* FP32 on wave-shared interaction lists: a node is accepted for
                                      64 bodies at once when the opening test holds at their
                                      bounding box (so for each body), offsets from FP64-formed
                                      group-relative positions, compensated accumulation; not
                                      available to the adjoint, whose walks are taped per body */
};

typedef struct {
    double G;          /* gravitational constant */
    double softening;  /* Plummer softening length (eps; eps^2 is added to r^2) */
    double theta;      /* opening angle: accept a cell when (2h)^2/d^2 < theta^2 */
    double dt;         /* substep length */
    int precision;     /* COSMOSIM_PREC_* */
    int device;        /* HIP device ordinal */
} CosmosimExtParams;

typedef struct {
    int n;                 /* bodies held by the handle */
    int n_live;            /* bodies with mass > 0 */
    int n_nodes;           /* nodes of the last tree (reference-layout count) */
    int max_depth;         /* deepest leaf of the last tree */
    int collision_runs;    /* groups of bodies sharing all 16 levels of the sorted key in the last build */
    int timed_walks;       /* walk launches covered by ms_walk */
    long long interactions;/* body-node interactions of the last walk */
    double root_center[3]; /* root cube of the last build */
    double root_half;
    double ms_build;       /* GPU time in build / walk / KDK since timing was enabled (0 if off) */
    double ms_walk;
    double ms_kdk;
    double ms_neighbors;   /* GPU time of neighbour-query traversals since timing was enabled */
    long long neighbor_queries; /* queries covered by ms_neighbors */
    double ms_sph;         /* GPU time of SPH density + force passes since timing was enabled */
    int list_fallbacks;    /* interaction-list walk wavefronts that fell back to per-body
                              traversal since timing was enabled */
    int reserved;
} CosmosimExtStats;

/* Seeded initial conditions (own generator; the reference's generators are
 * time-seeded).  recipe 0: single exponential disk + central SMBH; recipe 1:
 * two-disk merger.  orbit (may be NULL for the default) = relative position
 * (3) and relative velocity (3) of disk 2 with respect to disk 1.  n counts all
 * bodies including the SMBHs.  out must hold n records. */
enum { COSMOSIM_IC_DISK = 0, COSMOSIM_IC_MERGER = 1 };
int cosmosim_ext_generate(int recipe, int n, unsigned long long seed,
                          const double orbit[6], Body *out);
/* Default orbit of the merger recipe: separation 30, impact parameter 2,
 * relative speed 0.3 (written to orbit[6]). */
int cosmosim_ext_default_orbit(double orbit[6]);
/* d(body state)/d(orbit): for body i, x_i = local_i + coef[i]*r_rel and
 * v_i = vlocal_i + coef[i]*v_rel; writes coef[n] (disk 1: -M2/M, disk 2: M1/M). */
int cosmosim_ext_orbit_coefficients(int recipe, int n, double *coef);

/* Create a device-resident simulation of caller-supplied bodies (copied) and
 * compute the initial accelerations. */
int cosmosim_ext_create(const Body *bodies, int n, const CosmosimExtParams *params,
                        SimHandle *out);
int cosmosim_ext_set_params(SimHandle h, const CosmosimExtParams *params);
int cosmosim_ext_get_params(SimHandle h, CosmosimExtParams *params);
/* Advance k substeps on the device.  The host Body[] is NOT refreshed. */
int cosmosim_ext_step(SimHandle h, int substeps);
/* Copy the device state into the host Body[] returned by cosmosim_get_bodies. */
int cosmosim_ext_sync(SimHandle h);
/* Replace the device state with the host Body[] (after the caller edited it). */
int cosmosim_ext_upload(SimHandle h);
/* Rebuild the tree and recompute accelerations at the current positions. */
int cosmosim_ext_compute_forces(SimHandle h);
/* Device state as flat arrays (any pointer may be NULL): pos/vel/acc are
 * [n][3] row-major, mass is [n]. */
int cosmosim_ext_get_state(SimHandle h, double *pos, double *vel, double *acc, double *mass);
int cosmosim_ext_set_state(SimHandle h, const double *pos, const double *vel,
                           const double *acc, const double *mass);
/* FP64 direct-sum accelerations (the judge) for the k bodies idx[0..k) against
 * all bodies, written to acc[k][3]. */
int cosmosim_ext_direct_forces(SimHandle h, const int *idx, int k, double *acc);
/* Softened total energy and momentum of the device state (O(N^2) FP64). */
int cosmosim_ext_energy(SimHandle h, double *kinetic, double *potential, double momentum[3]);
/* Per-body interaction counts of the last explicit force evaluation (creation,
 * cosmosim_ext_compute_forces, the legacy walks), in original body order.  The
 * substep loops keep only the total (CosmosimExtStats.interactions); after one,
 * this returns COSMOSIM_E_STATE instead of stale counts. */
int cosmosim_ext_interaction_counts(SimHandle h, int *counts, int n);
int cosmosim_ext_get_stats(SimHandle h, CosmosimExtStats *out);
/* Enable and reset per-phase GPU event timing (events on the core's stream,
 * no host waits inside the substep loop). */
int cosmosim_ext_set_timing(SimHandle h, int enabled);
/* Write the last tree in reference layout (root 0); capacity in nodes. */
int cosmosim_ext_export_tree(SimHandle h, OctreeNode *pool, int capacity, int *pool_size);
/* Batched sphere queries over the device tree: the GPU form of
 * octree_find_neighbors (src/octree.c:236-318) for many centres at once, as the
 * SPH passes issue them (src/sph.c:57,120).  centres [k][3], radius [k];
 * out [k][cap] receives the first counts[q] <= cap bodies (original indices) of
 * query q, in the reference's visiting order (children in slot order), so the
 * lists equal the reference's, cap included.  Builds the tree if the handle's
 * positions have none yet. */
int cosmosim_ext_find_neighbors(SimHandle h, const double *centres, const double *radius, int k,
                                int cap, int *out, int *counts);
/* SPH of the handle's gas bodies on the device (replace sph_compute_density,
 * sph_compute_forces and sph_apply_cooling, src/sph.h:11-20): density, pressure
 * and the adaptive smoothing length; pressure-gradient plus viscosity
 * accelerations added to a; implicit cooling of the internal energy.  Results
 * equal src/sph.c's bit for bit on the same bodies and tree.  The density and
 * force passes build the tree if the handle's positions have none yet. */
int cosmosim_ext_sph_density(SimHandle h);
int cosmosim_ext_sph_forces(SimHandle h);
int cosmosim_ext_sph_cooling(SimHandle h, double dt);
/* Device-resident substeps of integrator_step with SPH enabled
 * (src/integrator.c:12-57): kick + drift, build, gravity, SPH density and
 * forces, kick, cooling. */
int cosmosim_ext_step_sph(SimHandle h, int substeps);
/* Snapshots for the host consumers, built on the device from the device state
 * and returned with one copy: the live bodies (mass > 0) in index order,
 * converted to float.
 *   _ue: the Unreal plugin's FCosmosimBodyGPU records, 16 floats each (pos,
 *        mass, vel, type, internal energy, density, lifetime, luminosity, spin,
 *        0), at most max_bodies of them - replaces the body loop of
 *        FCosmosimPhysicsRunnable::SnapshotToBuffer (CosmosimSubsystem.cpp:12-43);
 *   _render: the renderer's upload record, 10 floats each (pos, mass, vel,
 *        type, internal energy, density) - replaces the upload loop of
 *        src/renderer.c:348-364.
 * *count receives the number of records.  In quasar mode the host bodies are
 * current, so they are uploaded first. */
int cosmosim_ext_snapshot_ue(SimHandle h, float *out, int max_bodies, int *count);
int cosmosim_ext_snapshot_render(SimHandle h, float *out, int *count);
/* Quasar physics on the device (replace quasar_step and quasar_compact,
 * src/quasar.h:27-35).  CosmosimQuasarConfig has the byte layout of the
 * reference's QuasarConfig (src/quasar.h:6-22), so a caller's QuasarConfig can
 * be passed by pointer; jet_count is updated.  The step may add bodies (lobes,
 * jets) up to the handle's capacity; cosmosim_get_count reports the new count.
 * Results equal src/quasar.c's bit for bit except the positions and velocities
 * of spawned and recycled jets, which go through sin/cos/log (last-bit
 * differences between the device's and the host's libm).  The random stream and
 * the jets' knot phase (file-scope statics in the reference, starting at 1 and
 * 0) belong to the handle: _quasar_state gets (set = 0) or sets (set = 1) them. */
typedef struct {
    double accretion_radius, swallow_radius, isco_radius, viscosity_alpha, jet_speed, jet_mass,
        jet_lifetime, feedback_strength, eta_eff, eddington_k, accretion_smoothing;
    int jet_cap, jet_count, max_bodies, jet_ring_count;
} CosmosimQuasarConfig;
/* Grow the handle's body capacity (device and host) to at least capacity
 * records, keeping its bodies: room for the bodies a quasar step appends. */
int cosmosim_ext_reserve(SimHandle h, int capacity);
int cosmosim_ext_quasar_step(SimHandle h, CosmosimQuasarConfig *cfg, double dt);
int cosmosim_ext_quasar_compact(SimHandle h, int *n);
int cosmosim_ext_quasar_state(SimHandle h, unsigned long long *rng, int *phase, int set);
/* Get (set = 0) or set (set = 1) the QuasarConfig that the quasar-mode
 * cosmosim_step passes to its quasar steps (cosmosim_create fills it as
 * src/cosmosim_api.c:86-94 does; jet_count is updated by every step). */
int cosmosim_ext_quasar_config(SimHandle h, CosmosimQuasarConfig *cfg, int set);
/* Test hook for the capacity error path: the nth tree build from now (1 = the
 * next) fails as a tree needing more than the node capacity would, so a call
 * that runs through it must return COSMOSIM_E_CAPACITY; 0 cancels. */
int cosmosim_ext_debug_fail_build(SimHandle h, int nth);
/* Test hook: the last tree's body orders (csg_debug_orders): sorted order, the
 * FP32 walk's Peano-Hilbert wavefront order and the sorted keys, n_live each. */
int cosmosim_ext_debug_orders(SimHandle h, int *sorted_idx, int *wave_idx,
                              unsigned long long *keys, int n, int *n_live);
/* Last error recorded on the handle by a void legacy call (cosmosim_step). */
int cosmosim_ext_handle_error(SimHandle h);
/* Last error recorded on the calling thread by octree_* / integrator_*. */
int cosmosim_ext_last_error(void);
void cosmosim_ext_clear_error(void);
const char *cosmosim_ext_error_string(int code);
/* Adjoint of the force walk and of the leapfrog (no counterpart in the
 * reference).  Trees and opening decisions are frozen from the forward pass;
 * body masses are constants.  All arrays are [n][3] in body order.
 *  force_vjp:     xbar = (d a / d x)^T w at the current positions (rebuilds the tree).
 *  forces_frozen: forces at pos through the tree/decisions of the last force_vjp.
 *  direct_vjp:    the same VJP for the FP64 direct sum.
 *  record:        advance `substeps` from the current state, checkpointing every
 *                 `checkpoint_every`; keep_trees also keeps each tree for replay.
 *  backward:      gradients of a loss L(x_k, v_k) w.r.t. the recorded initial state;
 *                 mode 0 = recompute from checkpoints, 1 = time-reversed leapfrog;
 *                 stats (may be NULL) = {max reversed-state deviation, adjoint bytes}.
 *  replay:        frozen-tree forward of the recorded run from (pos0, vel0). */
int cosmosim_ext_force_vjp(SimHandle h, const double *w, double *xbar);
int cosmosim_ext_forces_frozen(SimHandle h, const double *pos, double *acc);
int cosmosim_ext_direct_vjp(SimHandle h, const double *w, double *xbar);
int cosmosim_ext_record(SimHandle h, int substeps, int checkpoint_every, int keep_trees);
int cosmosim_ext_backward(SimHandle h, const double *gpos, const double *gvel, int mode,
                          double *pos0_bar, double *vel0_bar, double stats[2]);
int cosmosim_ext_replay(SimHandle h, const double *pos0, const double *vel0, double *pos,
                        double *vel);

/* Allocation-free force evaluation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory while they run (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The variant of cosmosim_ext_compute_forces (tree build + walk at the given
 * positions) on caller-owned DEVICE arrays: x, y, z, m [n] in, ax, ay, az [n]
 * out (SoA, FP64), accelerations of bodies with m <= 0 set to 0.  All scratch
 * lives in one device array inside the library (about 19 MB per GPU, held
 * until the process ends), so n is at most 16384 (COSMOSIM_E_CAPACITY beyond).
 * Work is queued on `stream` (a hipStream_t; NULL = the default stream) and
 * the call returns without waiting; *status (a device int) receives 0, or
 * COSMOSIM_E_CAPACITY / COSMOSIM_E_HIP when the tree did not fit or the sort
 * stalled.  Calls that overlap on one GPU (other threads or streams) share the
 * scratch and race: order them on one stream. */
int cosmosim_ext_forces_noalloc(const double *x, const double *y, const double *z,
                                const double *m, int n, double G, double softening,
                                double theta, int precision, double *ax, double *ay,
                                double *az, int *status, void *stream);
/* Name and compute-unit count of the HIP device (name buffer of len bytes). */
int cosmosim_ext_device_info(int device, char *name, int len, int *compute_units);
/* Free and total device memory in bytes. */
int cosmosim_ext_device_memory(int device, unsigned long long *free_bytes,
                               unsigned long long *total_bytes);

#ifdef __cplusplus
}
#endif

#endif
