/* host_common.h - shared pieces of the C host layer (not installed). */
#ifndef COSMOSIM_HOST_COMMON_H
#define COSMOSIM_HOST_COMMON_H

#include <stdint.h>

#include "body.h"
#include "cosmosim_ext.h"
#include "csg.h"
#include "octree.h"

#define COSMOSIM_HIDDEN __attribute__((visibility("hidden")))

/* Thread-local record of the last error raised by a void legacy entry point. */
COSMOSIM_HIDDEN void cs_set_error(int code, const char *where);

/* The calling thread's device context for the legacy octree_* / integrator_*
 * functions (created on first use on device COSMOSIM_DEVICE, default 0). */
COSMOSIM_HIDDEN csg_ctx *cs_thread_ctx(void);
COSMOSIM_HIDDEN int cs_default_device(void);

/* Host <-> Body[] field transfer of the hot fields only (x..az), so that the
 * caller's other fields are never rewritten by the legacy path. */
COSMOSIM_HIDDEN int cs_download_kinematics(csg_ctx *ctx, Body *bodies, int n, int pos, int vel,
                                           int acc);

/* Fingerprint of the positions and masses of bodies[0..n). */
COSMOSIM_HIDDEN uint64_t cs_body_hash(const Body *bodies, int n);

/* Remember / test that `pool` holds the device tree of `bodies` (this thread). */
COSMOSIM_HIDDEN void cs_note_tree(const OctreeNode *pool, int pool_size, const Body *bodies, int n);
COSMOSIM_HIDDEN int cs_tree_matches(const OctreeNode *pool, int root, const Body *bodies, int n);
COSMOSIM_HIDDEN void cs_forget_tree(void);

/* Number of nodes reachable from root in a reference-layout pool (max index + 1). */
COSMOSIM_HIDDEN int cs_pool_extent(const OctreeNode *pool, int root);

/* One legacy substep (src/integrator.c:12-57) on ctx with host Body[] as the
 * source of truth; pool receives the tree (capacity pool_cap nodes). */
COSMOSIM_HIDDEN int cs_legacy_step(csg_ctx *ctx, Body *bodies, int n, double dt, double G,
                                   double softening, double theta, OctreeNode *pool, int pool_cap,
                                   int sph_enabled);
COSMOSIM_HIDDEN int cs_legacy_init(csg_ctx *ctx, Body *bodies, int n, double G, double softening,
                                   double theta, OctreeNode *pool, int pool_cap);

#endif
