/*
 * integrator.h - kick-drift-kick leapfrog entry points of the gravity core.
 *
 * Replaces: src/integrator.h:7-17.  Semantics follow src/integrator.c:4-57:
 *   live bodies (mass > 0):  v += (0.5*a)*dt;  x += v*dt;
 *   rebuild tree; recompute a;  [SPH hooks when sph_enabled];
 *   live bodies:             v += (0.5*a)*dt;  [SPH cooling when sph_enabled]
 * The kicks, drift, build, walk and, with sph_enabled, the SPH passes
 * (sph_compute_density, sph_compute_forces, sph_apply_cooling, src/sph.c) run
 * on the GPU where the reference calls them; the caller's pool receives the
 * freshly built tree and the body array every field, gas fields included.
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
