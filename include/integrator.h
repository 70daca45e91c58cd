/*
 * integrator.h - kick-drift-kick leapfrog entry points of the gravity core.
 *
 * Replaces: src/integrator.h:7-17.  Semantics follow src/integrator.c:4-57:
 *   live bodies (mass > 0):  v += (0.5*a)*dt;  x += v*dt;
 *   rebuild tree; recompute a;  [SPH hooks when sph_enabled];
 *   live bodies:             v += (0.5*a)*dt;  [SPH cooling when sph_enabled]
 * The kicks, drift, build and walk run on the GPU.  The SPH hooks
 * (sph_compute_density, sph_compute_forces, sph_apply_cooling) are the
 * caller's host code and are called exactly where the reference calls them,
 * with the caller's pool holding the freshly built tree.
 */
#ifndef INTEGRATOR_H
#define INTEGRATOR_H

#include "body.h"
#include "octree.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/integrator.c:12-57. */
void integrator_step(Body *bodies,
                     int n,
                     double dt,
                     double G,
                     double softening,
                     double theta,
                     OctreeNode *pool,
                     int sph_enabled);

/* Replaces src/integrator.c:4-10: one build + walk to seed a. */
void integrator_init_accelerations(
    Body *bodies, int n, double G, double softening, double theta, OctreeNode *pool);

#ifdef __cplusplus
}
#endif

#endif
