/*
 * body.h - per-body record shared by the host application and the gravity core.
 *
 * Replaces: src/body.h:4-27 (BodyType, Body).  The field names, order and the
 * resulting byte layout are the ABI that main.c, quasar.c, sph.c and the Unreal
 * plugin (ThirdParty/libcosmosim/include/body.h) compile against, so they are
 * kept identical; the static assertions below pin the x86-64 layout (168 bytes).
 *
 * The gravity core reads x, y, z and mass and writes x..az (bytes 0-79); every
 * other field is carried through untouched.  A body with mass <= 0 is "dead":
 * it is neither inserted into the tree, nor kicked, nor drifted, and its
 * acceleration is set to zero.
 */
#ifndef BODY_H
#define BODY_H

typedef enum {
    BODY_STAR = 0,
    BODY_GAS = 1,
    BODY_SMBH = 2,
    BODY_JET = 3,
    BODY_DUST = 4,
    BODY_LOBE = 5
} BodyType;

typedef struct {
    double x, y, z;          /* position (FP64) */
    double vx, vy, vz;       /* velocity */
    double ax, ay, az;       /* acceleration, written by the force walk */
    double mass;             /* <= 0 marks a dead body */
    BodyType type;
    double spin_x, spin_y, spin_z;
    double accretion_rate;
    double luminosity;
    double lifetime;
    double density;
    double pressure;
    double internal_energy;
    double smoothing_h;
} Body;

#if defined(__cplusplus)
#define COSMOSIM_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define COSMOSIM_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
#include <stddef.h>
COSMOSIM_STATIC_ASSERT(sizeof(Body) == 168, "Body must stay 168 bytes");
COSMOSIM_STATIC_ASSERT(offsetof(Body, x) == 0, "Body.x offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, vx) == 24, "Body.vx offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, ax) == 48, "Body.ax offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, mass) == 72, "Body.mass offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, type) == 80, "Body.type offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, spin_x) == 88, "Body.spin_x offset");
COSMOSIM_STATIC_ASSERT(offsetof(Body, smoothing_h) == 160, "Body.smoothing_h offset");

#endif
