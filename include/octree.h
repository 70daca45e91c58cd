/*
 * octree.h - tree node record and the tree entry points of the gravity core.
 *
 * Replaces: src/octree.h:6-36.  OctreeNode keeps the reference's 104-byte layout
 * so that a caller-owned pool (capacity 8*n nodes, root at index 0) can be read
 * by sph.c and by tests exactly as before.
 *
 * In this library the tree is built and walked on the GPU (HIP, gfx950); these
 * host functions upload the bodies, run the device kernels and copy the results
 * back, so their observable contract is the reference's:
 *   - octree_build writes a complete reference-layout tree into pool[0..*pool_size)
 *     with the root at index 0 and children[q] indexed by octant code q;
 *   - octree_compute_forces zeroes a for every body and adds the Barnes-Hut
 *     monopole acceleration for every live body (mass > 0).
 * Device errors cannot be returned through these void signatures; they are
 * recorded and can be read with cosmosim_ext_last_error() (cosmosim_ext.h).
 */
#ifndef OCTREE_H
#define OCTREE_H

#include "body.h"

typedef struct {
    double cx, cy, cz;                   /* centre of mass */
    double total_mass;
    double center_x, center_y, center_z; /* geometric centre of the cell */
    double half_size;
    int body_index;                      /* leaf: body index; internal/empty: -1 */
    int children[8];                     /* octant q = (x>=c) | (y<c)<<1 | (z<c)<<2; -1 if absent */
} OctreeNode;

COSMOSIM_STATIC_ASSERT(sizeof(OctreeNode) == 104, "OctreeNode must stay 104 bytes");
COSMOSIM_STATIC_ASSERT(offsetof(OctreeNode, total_mass) == 24, "OctreeNode.total_mass offset");
COSMOSIM_STATIC_ASSERT(offsetof(OctreeNode, center_x) == 32, "OctreeNode.center_x offset");
COSMOSIM_STATIC_ASSERT(offsetof(OctreeNode, half_size) == 56, "OctreeNode.half_size offset");
COSMOSIM_STATIC_ASSERT(offsetof(OctreeNode, body_index) == 64, "OctreeNode.body_index offset");
COSMOSIM_STATIC_ASSERT(offsetof(OctreeNode, children) == 68, "OctreeNode.children offset");

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/octree.c:114-169 (octree_build).  pool must hold >= 8*n nodes. */
void octree_build(OctreeNode *pool, int *pool_size, const Body *bodies, int n);

/* Replaces src/octree.c:220-234 (octree_compute_forces / compute_force_on_body
 * at :171-218).  When pool is the tree most recently written by octree_build on
 * this thread and root == 0, the walk runs on the device copy of that tree;
 * otherwise the caller's pool is uploaded and walked as given. */
void octree_compute_forces(const OctreeNode *pool,
                           int root,
                           Body *bodies,
                           int n,
                           double G,
                           double softening,
                           double theta);

/* Replaces src/octree.c:236-318 (octree_find_neighbors): sphere query over a
 * reference-layout pool, children visited in slot order, result capped at
 * buf_capacity.  Consumer: sph.c:57,120. */
int octree_find_neighbors(const OctreeNode *pool,
                          int root,
                          const Body *bodies,
                          int n,
                          double cx,
                          double cy,
                          double cz,
                          double radius,
                          int *neighbor_buf,
                          int buf_capacity);

#ifdef __cplusplus
}
#endif

#endif
