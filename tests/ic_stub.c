/*
 * ic_stub.c - TEST-ONLY initial-condition generators with the signatures of the
 * reference's src/initial_conditions.h (generate_spiral_galaxy, generate_merger,
 * generate_quasar_galaxy, generate_quasar_merger, generate_merger_dust), linked
 * into a test variant of libcosmosim.so (tests/_lib/libcosmosim_ic_stub.so) so
 * that cosmosim_create / cosmosim_step can be tested in every mode, quasar mode
 * included, without the caller's own generators.
 *
 * The recipes are this file's own and deterministic (fixed seeds): a uniform
 * disk with circular velocities around a central body; the quasar variants
 * make the central body a spinning, accreting black hole and the inner fifth
 * of the disk gas with SPH fields; the dust variant fills a bridge between the
 * two galaxies.  Bodies are fully written (the caller's array may hold stale
 * records).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "body.h"

static uint64_t stub_rng = 0x9e3779b97f4a7c15ull;

static double stub_uniform(void)
{
    /* splitmix64 */
    uint64_t z = (stub_rng += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void stub_disk(Body *b, int n, double cx, double cy, double mass, double radius,
                      double vx, double vy, double central_frac, uint64_t seed)
{
    stub_rng = seed;
    memset(b, 0, sizeof(Body) * (size_t)(n > 0 ? n : 0));
    if (n <= 0)
        return;
    const double mc = mass * central_frac;
    b[0].x = cx;
    b[0].y = cy;
    b[0].vx = vx;
    b[0].vy = vy;
    b[0].mass = mc;
    const double ms = n > 1 ? mass / (double)(n - 1) : 0.0;
    for (int i = 1; i < n; ++i) {
        const double r = radius * (0.03 + 0.97 * sqrt(stub_uniform()));
        const double phi = 6.283185307179586 * stub_uniform();
        const double menc = mc + mass * (r / radius) * (r / radius);
        const double v = sqrt(menc / r) * (0.95 + 0.1 * stub_uniform());
        b[i].x = cx + r * cos(phi);
        b[i].y = cy + r * sin(phi);
        b[i].z = 0.02 * radius * (stub_uniform() - 0.5);
        b[i].vx = vx - v * sin(phi);
        b[i].vy = vy + v * cos(phi);
        b[i].vz = 0.01 * v * (stub_uniform() - 0.5);
        b[i].mass = ms;
        b[i].type = BODY_STAR;
    }
}

void generate_spiral_galaxy(Body *bodies, int n, double cx, double cy, double galaxy_mass,
                            double disk_radius, double vx_bulk, double vy_bulk)
{
    stub_disk(bodies, n, cx, cy, galaxy_mass, disk_radius, vx_bulk, vy_bulk, 0.01, 11);
}

void generate_merger(Body *bodies, int n, double separation, double approach_vel)
{
    const int n1 = n / 2;
    stub_disk(bodies, n1, -0.5 * separation, 0.0, 2.0 * n1, 0.15 * separation, approach_vel,
              0.3 * approach_vel, 0.01, 12);
    stub_disk(bodies + n1, n - n1, 0.5 * separation, 0.0, 1.4 * n1, 0.12 * separation,
              -approach_vel, -0.3 * approach_vel, 0.01, 13);
}

static void stub_quasar(Body *b, int n, double cx, double cy, double mass, double radius,
                        double vx, double vy, double frac, uint64_t seed)
{
    stub_disk(b, n, cx, cy, mass, radius, vx, vy, frac, seed);
    if (n <= 0)
        return;
    b[0].type = BODY_SMBH;
    b[0].spin_y = 0.6;
    b[0].spin_z = 0.8;
    b[0].accretion_rate = mass * 1e-3;
    b[0].luminosity = 0.1 * b[0].accretion_rate;
    const double rin = 0.2 * radius;
    for (int i = 1; i < n; ++i) {
        const double dx = b[i].x - cx, dy = b[i].y - cy;
        if (dx * dx + dy * dy < rin * rin) {
            const double v2 = (b[i].vx - vx) * (b[i].vx - vx) + (b[i].vy - vy) * (b[i].vy - vy);
            b[i].type = BODY_GAS;
            b[i].internal_energy = 0.75 * v2;
            b[i].smoothing_h = 2.0 * radius / sqrt((double)n);
        }
    }
}

void generate_quasar_galaxy(Body *bodies, int n, double cx, double cy, double galaxy_mass,
                            double disk_radius, double vx_bulk, double vy_bulk,
                            double smbh_mass_frac)
{
    stub_quasar(bodies, n, cx, cy, galaxy_mass, disk_radius, vx_bulk, vy_bulk, smbh_mass_frac,
                21);
}

void generate_quasar_merger(Body *bodies, int n, double separation, double approach_vel,
                            double smbh_mass_frac)
{
    const int n1 = n / 2;
    const double m1 = 2.0 * n1, m2 = 1.2 * m1;
    const double v = 0.2 * sqrt((m1 + m2) / separation);
    stub_quasar(bodies, n1, -0.5 * separation, 0.0, m1, 0.15 * separation, approach_vel, v,
                smbh_mass_frac, 22);
    stub_quasar(bodies + n1, n - n1, 0.5 * separation, 0.0, m2, 0.195 * separation,
                -approach_vel, -0.85 * v, smbh_mass_frac, 23);
}

void generate_merger_dust(Body *bodies, int start_idx, int n_dust, double separation)
{
    stub_rng = 31;
    for (int i = 0; i < n_dust; ++i) {
        Body *b = &bodies[start_idx + i];
        memset(b, 0, sizeof(Body));
        const double t = stub_uniform();
        b->x = separation * (t - 0.5);
        b->y = 0.05 * separation * (stub_uniform() - 0.5);
        b->z = 0.02 * separation * (stub_uniform() - 0.5);
        b->vx = 0.2 * (stub_uniform() - 0.5);
        b->vy = 0.2 * (stub_uniform() - 0.5);
        b->mass = 0.4 + 0.6 * stub_uniform();
        b->type = BODY_DUST;
    }
}
