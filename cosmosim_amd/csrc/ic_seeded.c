/*
 * ic_seeded.c - seeded initial conditions for the benchmark and test workloads.
 *
 * No counterpart in the reference: its generators (src/initial_conditions.c)
 * are seeded from time(NULL) (:49) and use other profiles, so runs cannot be
 * repeated.  Recipes (G = 1, scale length Rd = 1):
 *
 *  DISK   n bodies = 1 central SMBH (index 0, mass 0.05, at rest at the origin)
 *         + n-1 equal-mass stars of total mass 1.  Stars: cylindrical radius R
 *         with density proportional to R exp(-R) (exponential surface density)
 *         truncated at 5 Rd, uniform azimuth, z ~ N(0, 0.1 Rd).  Velocity =
 *         circular speed sqrt(G M(<R) / R), with M(<R) the SMBH plus the
 *         analytic enclosed mass of the truncated disk, in the +phi direction,
 *         plus an isotropic Gaussian dispersion of 0.05 v_circ per component.
 *  MERGER two such disks of nominal mass 1.0 and 1.2 with SMBHs of 5% of each
 *         disk mass; n-2 stars of equal mass 2.2/(n-2) split between the disks
 *         in proportion to mass; disk 2 inclined 50 degrees about the x axis.
 *         Bodies: [SMBH1, stars1..., SMBH2, stars2...].  The orbit is the
 *         relative position r and velocity v of disk 2 with respect to disk 1
 *         (default r = (sqrt(30^2-2^2), 2, 0), v = (-0.3, 0, 0): separation 30,
 *         impact parameter 2, relative speed 0.3); both disks are placed in the
 *         centre-of-mass frame, disk 1 at -(M2/M) r, disk 2 at +(M1/M) r.
 *
 * Random numbers: splitmix64-seeded xoshiro256**, one stream for the whole body
 * array, drawn in body order.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "cosmosim_ext.h"

typedef struct {
    uint64_t s[4];
} Rng;

static uint64_t splitmix(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(Rng *r, uint64_t seed)
{
    for (int k = 0; k < 4; ++k)
        r->s[k] = splitmix(&seed);
}

static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

static uint64_t rng_next(Rng *r)
{
    uint64_t *s = r->s;
    const uint64_t out = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return out;
}

/* uniform in (0, 1] */
static double rng_u(Rng *r) { return ((double)(rng_next(r) >> 11) + 1.0) * 0x1.0p-53; }

static double rng_gauss(Rng *r)
{
    const double u1 = rng_u(r), u2 = rng_u(r);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

#define RD_TRUNC 5.0
#define SIGMA_Z 0.1
#define DISP_FRAC 0.05
#define SMBH_FRAC 0.05

/* Fraction of a truncated exponential disk's mass inside cylindrical radius R. */
static double disk_mass_fraction(double R)
{
    if (R >= RD_TRUNC)
        return 1.0;
    const double norm = 1.0 - (1.0 + RD_TRUNC) * exp(-RD_TRUNC);
    return (1.0 - (1.0 + R) * exp(-R)) / norm;
}

/* One disk: out[0] = SMBH, out[1..nstars] = stars, in the disk's own frame,
 * then rotated by `incl` about x and shifted by (pos, vel). */
static void make_disk(Rng *r, Body *out, int nstars, double mstar, double msmbh, double incl,
                      const double pos[3], const double vel[3])
{
    const double mdisk = mstar * nstars;
    const double ci = cos(incl), si = sin(incl);
    memset(out, 0, sizeof(Body) * (size_t)(nstars + 1));
    out[0].mass = msmbh;
    out[0].type = BODY_SMBH;
    for (int k = 1; k <= nstars; ++k) {
        double R;
        do {
            R = -log(rng_u(r) * rng_u(r)); /* Gamma(2,1): density R exp(-R) */
        } while (R > RD_TRUNC);
        const double phi = 2.0 * M_PI * rng_u(r);
        const double zz = SIGMA_Z * rng_gauss(r);
        const double cp = cos(phi), sp = sin(phi);
        const double menc = msmbh + mdisk * disk_mass_fraction(R);
        const double vc = sqrt(menc / R);
        const double g0 = rng_gauss(r), g1 = rng_gauss(r), g2 = rng_gauss(r);
        double p[3] = {R * cp, R * sp, zz};
        double v[3] = {-vc * sp + DISP_FRAC * vc * g0, vc * cp + DISP_FRAC * vc * g1,
                       DISP_FRAC * vc * g2};
        Body *b = &out[k];
        b->x = p[0];
        b->y = p[1] * ci - p[2] * si;
        b->z = p[1] * si + p[2] * ci;
        b->vx = v[0];
        b->vy = v[1] * ci - v[2] * si;
        b->vz = v[1] * si + v[2] * ci;
        b->mass = mstar;
        b->type = BODY_STAR;
    }
    for (int k = 0; k <= nstars; ++k) {
        out[k].x += pos[0];
        out[k].y += pos[1];
        out[k].z += pos[2];
        out[k].vx += vel[0];
        out[k].vy += vel[1];
        out[k].vz += vel[2];
    }
}

static void merger_split(int n, int *n1, int *n2, double *mstar)
{
    const int ns = n - 2;
    *mstar = 2.2 / ns;
    int a = (int)floor(1.0 / *mstar + 0.5);
    if (a < 1)
        a = 1;
    if (a > ns - 1)
        a = ns - 1;
    *n1 = a;
    *n2 = ns - a;
}

int cosmosim_ext_default_orbit(double orbit[6])
{
    if (!orbit)
        return COSMOSIM_E_ARG;
    orbit[0] = sqrt(30.0 * 30.0 - 2.0 * 2.0);
    orbit[1] = 2.0;
    orbit[2] = 0.0;
    orbit[3] = -0.3;
    orbit[4] = 0.0;
    orbit[5] = 0.0;
    return COSMOSIM_OK;
}

int cosmosim_ext_generate(int recipe, int n, unsigned long long seed, const double orbit[6],
                          Body *out)
{
    if (!out)
        return COSMOSIM_E_ARG;
    Rng r;
    rng_seed(&r, seed);
    const double zero[3] = {0.0, 0.0, 0.0};
    if (recipe == COSMOSIM_IC_DISK) {
        if (n < 2)
            return COSMOSIM_E_ARG;
        make_disk(&r, out, n - 1, 1.0 / (n - 1), SMBH_FRAC * 1.0, 0.0, zero, zero);
        return COSMOSIM_OK;
    }
    if (recipe == COSMOSIM_IC_MERGER) {
        if (n < 4)
            return COSMOSIM_E_ARG;
        double orb[6];
        if (orbit)
            memcpy(orb, orbit, sizeof(orb));
        else
            cosmosim_ext_default_orbit(orb);
        int n1, n2;
        double ms;
        merger_split(n, &n1, &n2, &ms);
        const double mb1 = SMBH_FRAC * 1.0, mb2 = SMBH_FRAC * 1.2;
        const double M1 = n1 * ms + mb1, M2 = n2 * ms + mb2, M = M1 + M2;
        const double p1[3] = {-M2 / M * orb[0], -M2 / M * orb[1], -M2 / M * orb[2]};
        const double v1[3] = {-M2 / M * orb[3], -M2 / M * orb[4], -M2 / M * orb[5]};
        const double p2[3] = {M1 / M * orb[0], M1 / M * orb[1], M1 / M * orb[2]};
        const double v2[3] = {M1 / M * orb[3], M1 / M * orb[4], M1 / M * orb[5]};
        make_disk(&r, out, n1, ms, mb1, 0.0, p1, v1);
        make_disk(&r, out + n1 + 1, n2, ms, mb2, 50.0 * M_PI / 180.0, p2, v2);
        return COSMOSIM_OK;
    }
    return COSMOSIM_E_ARG;
}

int cosmosim_ext_orbit_coefficients(int recipe, int n, double *coef)
{
    if (!coef || recipe != COSMOSIM_IC_MERGER || n < 4)
        return COSMOSIM_E_ARG;
    int n1, n2;
    double ms;
    merger_split(n, &n1, &n2, &ms);
    const double M1 = n1 * ms + SMBH_FRAC * 1.0, M2 = n2 * ms + SMBH_FRAC * 1.2, M = M1 + M2;
    for (int i = 0; i < n; ++i)
        coef[i] = (i <= n1) ? -M2 / M : M1 / M;
    return COSMOSIM_OK;
}
