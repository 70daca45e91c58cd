/*
 * api.c - the public handle API (src/cosmosim_api.h:35-42) and the extension
 * API (include/cosmosim_ext.h) over the HIP core.
 *
 * Handle state follows src/cosmosim_api.c:19-28 (bodies, pool, n_alloc,
 * current_n, config, sim_time, compact counter) plus a device context.  Two
 * stepping modes:
 *   quasar == 0 : device-resident.  The substep loop (kick+drift, build, walk,
 *                 kick) never leaves the GPU; the host Body[] is refreshed once
 *                 at the end of cosmosim_step (src/cosmosim_api.c:115-144).
 *   quasar == 1 : device-resident as well: each substep is the SPH-enabled
 *                 integrator step followed by the quasar step (src/quasar.c:477-510),
 *                 both on the device, which change n, masses, positions and
 *                 accelerations; compaction every 60 calls (src/cosmosim_api.c:
 *                 129-139) also runs there.  The host Body[] is refreshed once
 *                 at the end of cosmosim_step.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cosmosim_api.h"
#include "host_common.h"

#define API_G 1.0         /* src/cosmosim_api.c:10 */
#define API_SOFTENING 1.5 /* src/cosmosim_api.c:11 */
#define API_GALAXY_MASS 100.0
#define API_DISK_RADIUS 15.0
#define API_SEPARATION 40.0
#define API_APPROACH_VEL 0.15

/* The reference's QuasarConfig (src/quasar.h:6-22) has the layout of
 * csg_quasar_config; quasar_default_config's values (src/quasar.c:38-56). */
typedef csg_quasar_config QuasarConfigABI;

static QuasarConfigABI quasar_defaults(void)
{
    QuasarConfigABI q;
    memset(&q, 0, sizeof(q));
    q.accretion_radius = 3.0;
    q.swallow_radius = 0.3;
    q.isco_radius = 0.5;
    q.viscosity_alpha = 0.05;
    q.jet_speed = 20.0;
    q.jet_mass = 0.5;
    q.jet_lifetime = 5.0;
    q.feedback_strength = 1.0;
    q.eta_eff = 0.1;
    q.eddington_k = 0.5;
    q.accretion_smoothing = 0.95;
    q.jet_cap = 2000;
    q.jet_ring_count = 4;
    return q;
}

/* Initial conditions are the caller's code (src/initial_conditions.c), linked
 * into this library by the build's --ic-source option as the reference's
 * shared-library target links them: weak references, NULL when absent. */
extern void generate_spiral_galaxy(Body *, int, double, double, double, double, double, double)
    __attribute__((weak));
extern void generate_merger(Body *, int, double, double) __attribute__((weak));
extern void generate_quasar_galaxy(Body *, int, double, double, double, double, double, double,
                                   double) __attribute__((weak));
extern void generate_quasar_merger(Body *, int, double, double, double) __attribute__((weak));
extern void generate_merger_dust(Body *, int, int, double) __attribute__((weak));

typedef struct {
    Body *bodies;     /* pinned host array of n_alloc records */
    OctreeNode *pool; /* host pool, quasar mode only */
    int n_alloc;
    int current_n;
    CosmosimConfig config;
    QuasarConfigABI qcfg;
    double sim_time;
    int compact_counter;
    csg_ctx *ctx;
    CosmosimExtParams params;
    int host_stale; /* device state newer than bodies[] */
    int error;
} SimState;

static void state_free(SimState *s)
{
    if (!s)
        return;
    csg_destroy(s->ctx);
    csg_host_free(s->bodies);
    free(s->pool);
    free(s);
}

COSMOSIM_API CosmosimConfig cosmosim_default_config(void)
{
    CosmosimConfig cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.n_bodies = 20000;
    cfg.merger = 0;
    cfg.quasar = 0;
    cfg.dt = 0.005;
    cfg.theta = 0.5;
    cfg.smbh_mass_frac = 0.05;
    cfg.accretion_radius = 6.0;
    cfg.jet_speed = 15.0;
    cfg.feedback_strength = 0.3;
    cfg.substeps = 2;
    return cfg;
}

static void default_params(CosmosimExtParams *p, const CosmosimConfig *cfg)
{
    p->G = API_G;
    p->softening = API_SOFTENING;
    p->theta = cfg->theta;
    p->dt = cfg->dt;
    p->precision = COSMOSIM_PREC_FP64_STRICT;
    p->device = cs_default_device();
}

COSMOSIM_API SimHandle cosmosim_create(CosmosimConfig config)
{
    int n = config.n_bodies;
    if (n < 2)
        n = 2;
    const int n_dust = (config.merger && config.quasar) ? n / 5 : 0;
    const int n_alloc = config.quasar ? n + n_dust + n / 4 : n;

    int ndev = 0;
    if (csg_device_count(&ndev) != COSMOSIM_OK || cs_default_device() >= ndev) {
        cs_set_error(COSMOSIM_E_NO_DEVICE, "cosmosim_create");
        return NULL;
    }
    SimState *s = (SimState *)calloc(1, sizeof(SimState));
    if (!s)
        return NULL;
    s->bodies = (Body *)csg_host_alloc(sizeof(Body) * (size_t)n_alloc);
    if (!s->bodies) {
        state_free(s);
        return NULL;
    }
    s->n_alloc = n_alloc;
    s->config = config;
    default_params(&s->params, &config);

    const int have_ics = config.quasar
                             ? (generate_quasar_merger && generate_quasar_galaxy &&
                                generate_merger_dust)
                             : (config.merger ? generate_merger != NULL
                                              : generate_spiral_galaxy != NULL);
    if (!have_ics) {
        cs_set_error(COSMOSIM_E_UNSUPPORTED,
                     "cosmosim_create (the library was built without the caller's "
                     "initial_conditions.c: build with --ic-source)");
        state_free(s);
        return NULL;
    }
    if (config.quasar) {
        s->pool = (OctreeNode *)malloc(8 * (size_t)n_alloc * sizeof(OctreeNode));
        if (!s->pool) {
            state_free(s);
            return NULL;
        }
        if (config.merger) {
            generate_quasar_merger(s->bodies, n, API_SEPARATION, API_APPROACH_VEL,
                                   config.smbh_mass_frac);
            if (n_dust > 0)
                generate_merger_dust(s->bodies, n, n_dust, API_SEPARATION);
        } else {
            generate_quasar_galaxy(s->bodies, n, 0.0, 0.0, (double)n * 2.0, API_DISK_RADIUS, 0.0,
                                   0.0, config.smbh_mass_frac);
        }
    } else if (config.merger) {
        generate_merger(s->bodies, n, API_SEPARATION, API_APPROACH_VEL);
    } else {
        generate_spiral_galaxy(s->bodies, n, 0.0, 0.0, API_GALAXY_MASS, API_DISK_RADIUS, 0.0, 0.0);
    }

    const int initial_n = n + n_dust;
    if (csg_create(s->params.device, n_alloc, &s->ctx) != COSMOSIM_OK) {
        cs_set_error(COSMOSIM_E_NO_DEVICE, "cosmosim_create");
        state_free(s);
        return NULL;
    }
    int st;
    if (config.quasar) {
        st = cs_legacy_init(s->ctx, s->bodies, initial_n, API_G, API_SOFTENING, config.theta,
                            s->pool, 8 * n_alloc);
        QuasarConfigABI q = quasar_defaults();
        q.accretion_radius = config.accretion_radius;
        q.jet_speed = config.jet_speed;
        q.feedback_strength = config.feedback_strength;
        q.jet_cap = n / 10;
        q.max_bodies = n_alloc;
        s->qcfg = q;
    } else {
        st = csg_set_bodies(s->ctx, s->bodies, initial_n);
        if (!st)
            st = csg_build(s->ctx);
        if (!st)
            st = csg_walk(s->ctx, s->params.precision, API_G, API_SOFTENING, config.theta);
        if (!st)
            st = csg_get_bodies(s->ctx, s->bodies, initial_n);
    }
    if (st != COSMOSIM_OK) {
        cs_set_error(st, "cosmosim_create");
        state_free(s);
        return NULL;
    }
    s->current_n = initial_n;
    s->sim_time = 0.0;
    s->compact_counter = 0;
    return (SimHandle)s;
}

static int sync_host(SimState *s)
{
    if (!s->host_stale)
        return COSMOSIM_OK;
    int st = csg_get_bodies(s->ctx, s->bodies, s->current_n);
    if (st == COSMOSIM_OK)
        s->host_stale = 0;
    return st;
}

COSMOSIM_API void cosmosim_step(SimHandle handle)
{
    if (!handle)
        return;
    SimState *s = (SimState *)handle;
    int st = COSMOSIM_OK;
    if (s->config.quasar) {
        /* integrator_step with SPH, then quasar_step, per substep; all on the
         * device (src/cosmosim_api.c:124-139) */
        for (int sub = 0; sub < s->config.substeps && st == COSMOSIM_OK; ++sub) {
            st = csg_substeps_sph(s->ctx, 1, s->config.dt, API_G, API_SOFTENING,
                                  s->config.theta, COSMOSIM_PREC_FP64_STRICT);
            if (st == COSMOSIM_OK)
                st = csg_quasar_step(s->ctx, &s->qcfg, s->config.dt);
        }
        s->compact_counter++;
        if (st == COSMOSIM_OK && s->compact_counter >= 60) {
            st = csg_quasar_compact(s->ctx, NULL);
            s->compact_counter = 0;
        }
        s->current_n = csg_size(s->ctx);
        if (st == COSMOSIM_OK) {
            s->host_stale = 1;
            st = sync_host(s);
        }
    } else {
        st = csg_substeps(s->ctx, s->config.substeps, s->config.dt, s->params.G,
                          s->params.softening, s->config.theta, s->params.precision);
        if (st == COSMOSIM_OK) {
            s->host_stale = 1;
            st = sync_host(s);
        }
    }
    s->error = st;
    if (st != COSMOSIM_OK)
        fprintf(stderr, "cosmosim: cosmosim_step failed: %s\n", cosmosim_ext_error_string(st));
    s->sim_time += s->config.dt * s->config.substeps;
}

COSMOSIM_API const Body *cosmosim_get_bodies(SimHandle handle)
{
    if (!handle)
        return NULL;
    SimState *s = (SimState *)handle;
    int st = sync_host(s);
    if (st != COSMOSIM_OK)
        s->error = st;
    return s->bodies;
}

COSMOSIM_API int cosmosim_get_count(SimHandle handle)
{
    if (!handle)
        return 0;
    return ((SimState *)handle)->current_n;
}

COSMOSIM_API int cosmosim_get_active_count(SimHandle handle)
{
    if (!handle)
        return 0;
    const SimState *s = (const SimState *)handle;
    int count = 0;
    for (int i = 0; i < s->current_n; i++)
        if (s->bodies[i].mass > 0.0)
            count++;
    return count;
}

COSMOSIM_API double cosmosim_get_sim_time(SimHandle handle)
{
    if (!handle)
        return 0.0;
    return ((SimState *)handle)->sim_time;
}

COSMOSIM_API void cosmosim_destroy(SimHandle handle)
{
    if (!handle)
        return;
    state_free((SimState *)handle);
}

/* =================================================================== ext */
COSMOSIM_API const char *cosmosim_ext_error_string(int code)
{
    switch (code) {
    case COSMOSIM_OK: return "ok";
    case COSMOSIM_E_ARG: return "invalid argument";
    case COSMOSIM_E_NO_DEVICE: return "no usable HIP device";
    case COSMOSIM_E_ALLOC: return "allocation failed";
    case COSMOSIM_E_HIP: return "HIP runtime error";
    case COSMOSIM_E_CAPACITY: return "capacity exceeded (tree nodes or output buffer)";
    case COSMOSIM_E_STATE: return "invalid state for this call";
    case COSMOSIM_E_UNSUPPORTED: return "unsupported in this build";
    default: return "unknown error";
    }
}

static int valid_params(const CosmosimExtParams *p)
{
    return p && p->precision >= COSMOSIM_PREC_FP64_STRICT && p->precision <= COSMOSIM_PREC_FP32_LIST &&
           p->theta >= 0.0 && p->softening >= 0.0 && p->device >= 0;
}

COSMOSIM_API int cosmosim_ext_create(const Body *bodies, int n, const CosmosimExtParams *params,
                                     SimHandle *out)
{
    if (!out || n < 0 || (n > 0 && !bodies) || !valid_params(params))
        return COSMOSIM_E_ARG;
    *out = NULL;
    int ndev = 0;
    if (csg_device_count(&ndev) != COSMOSIM_OK || params->device >= ndev)
        return COSMOSIM_E_NO_DEVICE;
    SimState *s = (SimState *)calloc(1, sizeof(SimState));
    if (!s)
        return COSMOSIM_E_ALLOC;
    s->bodies = (Body *)csg_host_alloc(sizeof(Body) * (size_t)(n > 0 ? n : 1));
    if (!s->bodies) {
        state_free(s);
        return COSMOSIM_E_ALLOC;
    }
    if (n > 0)
        memcpy(s->bodies, bodies, sizeof(Body) * (size_t)n);
    s->n_alloc = n;
    s->current_n = n;
    s->config = cosmosim_default_config();
    s->config.n_bodies = n;
    s->config.dt = params->dt;
    s->config.theta = params->theta;
    s->config.substeps = 1;
    s->params = *params;
    int st = csg_create(params->device, n, &s->ctx);
    if (!st)
        st = csg_set_bodies(s->ctx, s->bodies, n);
    if (!st)
        st = csg_build(s->ctx);
    if (!st)
        st = csg_walk(s->ctx, params->precision, params->G, params->softening, params->theta);
    if (!st)
        st = csg_get_bodies(s->ctx, s->bodies, n);
    if (st) {
        state_free(s);
        return st;
    }
    *out = (SimHandle)s;
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_set_params(SimHandle h, const CosmosimExtParams *p)
{
    if (!h || !valid_params(p))
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (p->device != s->params.device)
        return COSMOSIM_E_STATE;
    s->params = *p;
    s->config.dt = p->dt;
    s->config.theta = p->theta;
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_get_params(SimHandle h, CosmosimExtParams *p)
{
    if (!h || !p)
        return COSMOSIM_E_ARG;
    *p = ((SimState *)h)->params;
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_step(SimHandle h, int substeps)
{
    if (!h || substeps < 0)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (s->config.quasar)
        return COSMOSIM_E_STATE;
    int st = csg_substeps(s->ctx, substeps, s->params.dt, s->params.G, s->params.softening,
                          s->params.theta, s->params.precision);
    if (st == COSMOSIM_OK) {
        s->host_stale = 1;
        s->sim_time += s->params.dt * substeps;
    }
    return st;
}

COSMOSIM_API int cosmosim_ext_sync(SimHandle h)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (s->config.quasar)
        return COSMOSIM_OK;
    s->host_stale = 1;
    return sync_host(s);
}

COSMOSIM_API int cosmosim_ext_upload(SimHandle h)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    s->host_stale = 0;
    return csg_set_bodies(s->ctx, s->bodies, s->current_n);
}

COSMOSIM_API int cosmosim_ext_compute_forces(SimHandle h)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_build(s->ctx);
    if (!st)
        st = csg_walk(s->ctx, s->params.precision, s->params.G, s->params.softening,
                      s->params.theta);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_get_state(SimHandle h, double *pos, double *vel, double *acc,
                                        double *mass)
{
    if (!h)
        return COSMOSIM_E_ARG;
    return csg_get_state(((SimState *)h)->ctx, pos, vel, acc, mass);
}

COSMOSIM_API int cosmosim_ext_set_state(SimHandle h, const double *pos, const double *vel,
                                        const double *acc, const double *mass)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_set_state(s->ctx, pos, vel, acc, mass);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_direct_forces(SimHandle h, const int *idx, int k, double *acc)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    return csg_direct_forces(s->ctx, idx, k, s->params.G, s->params.softening, acc);
}

COSMOSIM_API int cosmosim_ext_energy(SimHandle h, double *kinetic, double *potential,
                                     double momentum[3])
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    return csg_energy(s->ctx, s->params.G, s->params.softening, kinetic, potential, momentum);
}

COSMOSIM_API int cosmosim_ext_interaction_counts(SimHandle h, int *counts, int n)
{
    if (!h)
        return COSMOSIM_E_ARG;
    return csg_interaction_counts(((SimState *)h)->ctx, counts, n);
}

COSMOSIM_API int cosmosim_ext_get_stats(SimHandle h, CosmosimExtStats *out)
{
    if (!h || !out)
        return COSMOSIM_E_ARG;
    csg_stats cs;
    int st = csg_get_stats(((SimState *)h)->ctx, &cs);
    if (st)
        return st;
    memset(out, 0, sizeof(*out));
    out->n = cs.n;
    out->n_live = cs.n_live;
    out->n_nodes = cs.n_nodes;
    out->max_depth = cs.max_depth;
    out->collision_runs = cs.collision_runs;
    out->interactions = cs.interactions;
    for (int a = 0; a < 3; ++a)
        out->root_center[a] = cs.root[a];
    out->root_half = cs.root[3];
    out->ms_build = cs.ms_build;
    out->ms_walk = cs.ms_walk;
    out->ms_kdk = cs.ms_kdk;
    out->timed_walks = cs.timed_walks;
    out->ms_neighbors = cs.ms_neighbors;
    out->neighbor_queries = cs.neighbor_queries;
    out->ms_sph = cs.ms_sph;
    out->list_fallbacks = cs.list_fallbacks;
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_set_timing(SimHandle h, int enabled)
{
    if (!h)
        return COSMOSIM_E_ARG;
    return csg_set_timing(((SimState *)h)->ctx, enabled);
}

COSMOSIM_API int cosmosim_ext_export_tree(SimHandle h, OctreeNode *pool, int capacity,
                                          int *pool_size)
{
    if (!h)
        return COSMOSIM_E_ARG;
    return csg_export_pool(((SimState *)h)->ctx, pool, capacity, pool_size);
}

COSMOSIM_API int cosmosim_ext_find_neighbors(SimHandle h, const double *centres,
                                             const double *radius, int k, int cap, int *out,
                                             int *counts)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_find_neighbors(s->ctx, centres, radius, k, cap, out, counts);
    if (st == COSMOSIM_E_STATE) {
        /* no tree for the current positions yet: build it, then query */
        st = csg_build(s->ctx);
        if (!st)
            st = csg_find_neighbors(s->ctx, centres, radius, k, cap, out, counts);
    }
    return st;
}

/* A pass that needs the tree builds it when the positions have none. */
static int with_tree(SimState *s, int (*pass)(csg_ctx *))
{
    int st = pass(s->ctx);
    if (st == COSMOSIM_E_STATE) {
        st = csg_build(s->ctx);
        if (!st)
            st = pass(s->ctx);
    }
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_sph_density(SimHandle h)
{
    return h ? with_tree((SimState *)h, csg_sph_density) : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_sph_forces(SimHandle h)
{
    return h ? with_tree((SimState *)h, csg_sph_forces) : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_sph_cooling(SimHandle h, double dt)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_sph_cooling(s->ctx, dt);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_step_sph(SimHandle h, int substeps)
{
    if (!h || substeps < 0)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (s->config.quasar)
        return COSMOSIM_E_STATE;
    int st = csg_substeps_sph(s->ctx, substeps, s->params.dt, s->params.G, s->params.softening,
                              s->params.theta, s->params.precision);
    if (st == COSMOSIM_OK) {
        s->host_stale = 1;
        s->sim_time += s->params.dt * substeps;
    }
    return st;
}

static int snapshot(SimHandle h, int fmt, float *out, int max_bodies, int *count)
{
    if (!h || !count)
        return COSMOSIM_E_ARG;
    return csg_snapshot(((SimState *)h)->ctx, fmt, out, max_bodies, count);
}

COSMOSIM_API int cosmosim_ext_reserve(SimHandle h, int capacity)
{
    if (!h || capacity < 0)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (capacity <= s->n_alloc)
        return COSMOSIM_OK;
    int st = sync_host(s);
    if (st)
        return st;
    Body *nb = (Body *)csg_host_alloc(sizeof(Body) * (size_t)capacity);
    if (!nb)
        return COSMOSIM_E_ALLOC;
    if (s->current_n > 0)
        memcpy(nb, s->bodies, sizeof(Body) * (size_t)s->current_n);
    if ((st = csg_reserve(s->ctx, capacity))) {
        csg_host_free(nb);
        return st;
    }
    csg_host_free(s->bodies);
    s->bodies = nb;
    s->n_alloc = capacity;
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_quasar_step(SimHandle h, CosmosimQuasarConfig *cfg, double dt)
{
    if (!h || !cfg)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_quasar_step(s->ctx, (csg_quasar_config *)cfg, dt);
    s->current_n = csg_size(s->ctx);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_quasar_compact(SimHandle h, int *n)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_quasar_compact(s->ctx, n);
    s->current_n = csg_size(s->ctx);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_quasar_state(SimHandle h, unsigned long long *rng, int *phase,
                                           int set)
{
    return h ? csg_quasar_state(((SimState *)h)->ctx, rng, phase, set) : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_quasar_config(SimHandle h, CosmosimQuasarConfig *cfg, int set)
{
    if (!h || !cfg)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (set)
        memcpy(&s->qcfg, cfg, sizeof(s->qcfg));
    else
        memcpy(cfg, &s->qcfg, sizeof(s->qcfg));
    return COSMOSIM_OK;
}

COSMOSIM_API int cosmosim_ext_snapshot_ue(SimHandle h, float *out, int max_bodies, int *count)
{
    return snapshot(h, 0, out, max_bodies, count);
}

COSMOSIM_API int cosmosim_ext_snapshot_render(SimHandle h, float *out, int *count)
{
    return snapshot(h, 1, out, h ? ((SimState *)h)->current_n : 0, count);
}

COSMOSIM_API int cosmosim_ext_debug_fail_build(SimHandle h, int nth)
{
    return h ? csg_debug_fail_build(((SimState *)h)->ctx, nth) : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_debug_orders(SimHandle h, int *sorted_idx, int *wave_idx,
                                           unsigned long long *keys, int n, int *n_live)
{
    return h ? csg_debug_orders(((SimState *)h)->ctx, sorted_idx, wave_idx, keys, n, n_live)
             : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_handle_error(SimHandle h)
{
    return h ? ((SimState *)h)->error : COSMOSIM_E_ARG;
}

COSMOSIM_API int cosmosim_ext_device_info(int device, char *name, int len, int *compute_units)
{
    return csg_device_info(device, name, len, compute_units);
}

COSMOSIM_API int cosmosim_ext_device_memory(int device, unsigned long long *free_bytes,
                                            unsigned long long *total_bytes)
{
    return csg_device_memory(device, free_bytes, total_bytes);
}

/* ---------------------------------------------------------------- adjoint */
COSMOSIM_API int cosmosim_ext_force_vjp(SimHandle h, const double *w, double *xbar)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    int st = csg_build(s->ctx);
    if (!st)
        st = csg_walk_taped(s->ctx, s->params.precision, s->params.G, s->params.softening,
                            s->params.theta);
    if (!st)
        st = csg_force_vjp(s->ctx, w, xbar);
    if (!st)
        s->host_stale = 1;
    return st;
}

COSMOSIM_API int cosmosim_ext_forces_frozen(SimHandle h, const double *pos, double *acc)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    s->host_stale = 1;
    return csg_forces_frozen(s->ctx, pos, acc);
}

COSMOSIM_API int cosmosim_ext_direct_vjp(SimHandle h, const double *w, double *xbar)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    return csg_direct_vjp(s->ctx, w, xbar, s->params.G, s->params.softening);
}

COSMOSIM_API int cosmosim_ext_record(SimHandle h, int substeps, int checkpoint_every,
                                     int keep_trees)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    if (s->config.quasar)
        return COSMOSIM_E_STATE;
    int st = csg_record(s->ctx, substeps, s->params.dt, s->params.G, s->params.softening,
                        s->params.theta, s->params.precision, checkpoint_every, keep_trees);
    if (st == COSMOSIM_OK) {
        s->host_stale = 1;
        s->sim_time += s->params.dt * substeps;
    }
    return st;
}

COSMOSIM_API int cosmosim_ext_backward(SimHandle h, const double *gpos, const double *gvel,
                                       int mode, double *pos0_bar, double *vel0_bar,
                                       double stats[2])
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    s->host_stale = 1;
    return csg_backward(s->ctx, gpos, gvel, mode, pos0_bar, vel0_bar, stats);
}

COSMOSIM_API int cosmosim_ext_replay(SimHandle h, const double *pos0, const double *vel0,
                                     double *pos, double *vel)
{
    if (!h)
        return COSMOSIM_E_ARG;
    SimState *s = (SimState *)h;
    s->host_stale = 1;
    return csg_replay(s->ctx, pos0, vel0, pos, vel);
}

