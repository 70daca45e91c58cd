/*
 * cosmosim_api.h - public opaque-handle C-ABI of libcosmosim.
 *
 * Replaces: src/cosmosim_api.h:1-48 (CosmosimConfig at :16-27, SimHandle at :29,
 * the eight exports at :35-42).  Signatures, struct layout (72 bytes) and the
 * error convention are unchanged so the Unreal plugin (CosmosimSubsystem.cpp,
 * CosmosimModule.cpp dlopen) links and runs against this library as before:
 *   - cosmosim_create returns NULL when any host or device allocation fails;
 *   - every entry point accepts a NULL handle (returns 0 / NULL / does nothing);
 *   - cosmosim_step leaves the host Body[] returned by cosmosim_get_bodies fully
 *     up to date before it returns;
 *   - G = 1 and softening = 1.5 are fixed inside create/step, as in
 *     src/cosmosim_api.c:10-11.
 * Device-side failures inside the void cosmosim_step are recorded and can be read
 * with cosmosim_ext_handle_error() from cosmosim_ext.h.
 */
#ifndef COSMOSIM_API_H
#define COSMOSIM_API_H

#include "body.h"

#ifdef _WIN32
#ifdef COSMOSIM_BUILDING_DLL
#define COSMOSIM_API __declspec(dllexport)
#else
#define COSMOSIM_API __declspec(dllimport)
#endif
#else
#define COSMOSIM_API __attribute__((visibility("default")))
#endif

/* Field order and types are the ABI; the meaning of each field:
 *   n_bodies        bodies the initial conditions make (at least 2)
 *   merger, quasar  flags: two galaxies instead of one; black-hole physics
 *                   (SPH, accretion, jets) between substeps
 *   dt, theta       substep length; the cell-acceptance angle of the walk
 *   smbh_mass_frac .. feedback_strength: quasar-mode black-hole parameters
 *   substeps        KDK substeps advanced by one cosmosim_step */
typedef struct {
    int n_bodies;
    int merger;
    int quasar;
    double dt;
    double theta;
    double smbh_mass_frac;
    double accretion_radius;
    double jet_speed;
    double feedback_strength;
    int substeps;
} CosmosimConfig;

COSMOSIM_STATIC_ASSERT(sizeof(CosmosimConfig) == 72, "CosmosimConfig must stay 72 bytes");
COSMOSIM_STATIC_ASSERT(offsetof(CosmosimConfig, dt) == 16, "CosmosimConfig.dt offset");
COSMOSIM_STATIC_ASSERT(offsetof(CosmosimConfig, substeps) == 64, "CosmosimConfig.substeps offset");

typedef void *SimHandle;

#ifdef __cplusplus
extern "C" {
#endif

/* src/cosmosim_api.c:30-45: n 20000, dt 0.005, theta 0.5, substeps 2, ... */
COSMOSIM_API CosmosimConfig cosmosim_default_config(void);
/* src/cosmosim_api.c:47-113 */
COSMOSIM_API SimHandle cosmosim_create(CosmosimConfig config);
/* src/cosmosim_api.c:115-144 */
COSMOSIM_API void cosmosim_step(SimHandle handle);
/* src/cosmosim_api.c:146-151: borrowed host pointer, valid until destroy */
COSMOSIM_API const Body *cosmosim_get_bodies(SimHandle handle);
/* src/cosmosim_api.c:153-158 */
COSMOSIM_API int cosmosim_get_count(SimHandle handle);
/* src/cosmosim_api.c:160-171 */
COSMOSIM_API int cosmosim_get_active_count(SimHandle handle);
/* src/cosmosim_api.c:173-178 */
COSMOSIM_API double cosmosim_get_sim_time(SimHandle handle);
/* src/cosmosim_api.c:180-188 */
COSMOSIM_API void cosmosim_destroy(SimHandle handle);

#ifdef __cplusplus
}
#endif

#endif /* COSMOSIM_API_H */
