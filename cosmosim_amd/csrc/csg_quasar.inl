// csg_quasar.inl - quasar physics on the device state: the GPU form of
// quasar_step and quasar_compact (src/quasar.c:11-523).  Included at the end of
// csg_core.hip.
//
// The reference's step is sequential in places; the device keeps its results:
//   * accretion (quasar.c:123-198): for each black hole in index order, bodies
//     are visited in index order and a swallowed body changes the black hole's
//     mass, velocity and spin for the bodies after it.  The device runs passes:
//     find the first body (from the current start) that the black hole swallows
//     under its current state (parallel, atomic minimum), apply drag and gas
//     conversion to the bodies before it (parallel, under that same state), then
//     one thread swallows it and moves the start past it.  Every body therefore
//     sees the black-hole state the reference's loop would show it.
//   * the random stream (xorshift64, quasar.c:11-36) is drawn in the reference's
//     order: lobes of dying jets in index order, recycled jets in index order,
//     then the spawned jets; one thread fills the draws of each phase and the
//     bodies take theirs by their rank (a block scan over index order).
//   * jet spawning (quasar.c:232-391) is a few dozen bodies per step and runs in
//     one thread in the reference's order; the knot phase is the context's.
// Expressions keep the reference's order; sqrt and divide are IEEE, so every
// part is bit-identical to the reference except the positions and velocities of
// spawned and recycled jets, which go through sin/cos/log (the device's libm
// and the host's may differ in the last bit).
//
// State: SoA positions, velocities, accelerations and masses; type, spin,
// accretion rate, luminosity, lifetime and the gas fields in the device body
// records (c->aos).  The step may append bodies (lobes, jets) up to the
// context's capacity.

struct QDev {                 // device-side step state
    double swallowed;         // mass swallowed this step, summed in event order
    unsigned long long rng;   // xorshift64 state
    int n;                    // body count (grows with lobes and jets)
    int jet_count;            // QuasarConfig.jet_count
    int i0;                   // accretion: first index of the current pass
    int first;                // accretion: first event index of the pass (n: none)
    int done;                 // accretion: no event left for the current black hole
    int nsmbh;                // black holes (type SMBH, any mass) in list[]
    int cnt;                  // bodies of the current phase that draw (lobes, recycled)
    int s_recycle;            // first live black hole (-1: none)
};

#define Q_MAX_SMBH 4096

__device__ __forceinline__ double q_uniform(unsigned long long &st) {
    unsigned long long x = st;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    st = x;
    return (x >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void q_kill(int i, double *m, double *vx, double *vy, double *vz,
                                       double *ax, double *ay, double *az) {
    m[i] = 0.0;
    vx[i] = vy[i] = vz[i] = 0.0;
    ax[i] = ay[i] = az[i] = 0.0;
}

struct QArrays {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m;
    Body *aos;
};

// ------------------------------------------------------------ black holes
__global__ __launch_bounds__(CSG_TPB) void k_q_list(QArrays a, int n, int *list, QDev *d) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n || a.aos[i].type != BODY_SMBH) return;
    const int k = atomicAdd(&d->nsmbh, 1);
    if (k < Q_MAX_SMBH) list[k] = i;
}

// One thread: the list in index order, the first live black hole.
__global__ void k_q_sort(QArrays a, int *list, QDev *d) {
    const int k = min(d->nsmbh, Q_MAX_SMBH);
    for (int p = 1; p < k; ++p) {
        const int v = list[p];
        int q = p - 1;
        while (q >= 0 && list[q] > v) {
            list[q + 1] = list[q];
            --q;
        }
        list[q + 1] = v;
    }
    d->s_recycle = -1;
    for (int p = 0; p < k; ++p)
        if (a.m[list[p]] > 0.0) {
            d->s_recycle = list[p];
            break;
        }
}

// ------------------------------------------------------------ accretion
// What black hole s does to body i under its current state: 0 nothing, 1 drag
// and possibly conversion, 2 swallowed (quasar.c:137-196).
__device__ __forceinline__ int q_acc_kind(const QArrays &a, int s, int i,
                                          const csg_quasar_config &q, double &r, double &rc) {
    if (i == s || a.m[i] <= 0.0 || a.aos[i].type == BODY_SMBH) return 0;
    const double dx = a.x[i] - a.x[s], dy = a.y[i] - a.y[s], dz = a.z[i] - a.z[s];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > q.accretion_radius * q.accretion_radius) return 0;
    if (d2 < q.swallow_radius * q.swallow_radius) return 2;
    r = sqrt(d2);
    const double ux = a.vx[i] - a.vx[s], uy = a.vy[i] - a.vy[s], uz = a.vz[i] - a.vz[s];
    const double lx = dy * uz - dz * uy, ly = dz * ux - dx * uz, lz = dx * uy - dy * ux;
    rc = (lx * lx + ly * ly + lz * lz) / (a.m[s] * r);
    const bool gas = a.aos[i].type == BODY_GAS || rc > q.isco_radius;
    return (gas && r < q.isco_radius) ? 2 : 1;
}

// Gas conversion (with the gas fields' initial values) and disk drag of body i
// (quasar.c:99-121, 171-190).
__device__ __forceinline__ void q_acc_effects(QArrays &a, int s, int i, double rc,
                                              const csg_quasar_config &q) {
    Body &b = a.aos[i];
    if (rc > q.isco_radius && b.type != BODY_GAS) {
        b.type = BODY_GAS;
        if (b.smoothing_h < 1e-10) b.smoothing_h = q.accretion_radius * 0.3;
        if (b.internal_energy < 1e-10)
            b.internal_energy =
                0.001 * (a.vx[i] * a.vx[i] + a.vy[i] * a.vy[i] + a.vz[i] * a.vz[i]) + 0.01;
    }
    if (b.type != BODY_GAS) return;
    const double rx = a.x[i] - a.x[s], ry = a.y[i] - a.y[s], rz = a.z[i] - a.z[s];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1e-15) return;
    const double ux = a.vx[i] - a.vx[s], uy = a.vy[i] - a.vy[s], uz = a.vz[i] - a.vz[s];
    const double vr = (ux * rx + uy * ry + uz * rz) / r;
    a.ax[i] -= q.viscosity_alpha * (vr * rx / r);
    a.ay[i] -= q.viscosity_alpha * (vr * ry / r);
    a.az[i] -= q.viscosity_alpha * (vr * rz / r);
}

__global__ void k_acc_begin(QArrays a, int s, QDev *d) {
    d->i0 = 0;
    d->first = d->n;
    d->done = a.m[s] > 0.0 ? 0 : 1;   // a dead black hole accretes nothing
    if (d->done) d->i0 = d->n;
}

__global__ __launch_bounds__(CSG_TPB) void k_acc_first(QArrays a, int s, csg_quasar_config q,
                                                       QDev *d) {
    const int i = d->i0 + (int)(blockIdx.x * CSG_TPB + threadIdx.x);
    if (d->done || i >= d->n) return;
    double r = 0.0, rc = 0.0;
    if (q_acc_kind(a, s, i, q, r, rc) == 2) atomicMin(&d->first, i);
}

__global__ __launch_bounds__(CSG_TPB) void k_acc_apply(QArrays a, int s, csg_quasar_config q,
                                                       QDev *d) {
    const int i = d->i0 + (int)(blockIdx.x * CSG_TPB + threadIdx.x);
    if (i >= d->first) return;
    double r = 0.0, rc = 0.0;
    if (q_acc_kind(a, s, i, q, r, rc) == 1) q_acc_effects(a, s, i, rc, q);
}

// One thread: the pass's event (conversion and drag on the ISCO path, then the
// swallow, quasar.c:58-97), the next pass's start.
__global__ void k_acc_event(QArrays a, int s, csg_quasar_config q, QDev *d) {
    if (d->done) return;
    const int j = d->first;
    if (j >= d->n) {
        d->done = 1;
        d->i0 = d->n;
        return;
    }
    double r = 0.0, rc = 0.0;
    const double d2x = a.x[j] - a.x[s], d2y = a.y[j] - a.y[s], d2z = a.z[j] - a.z[s];
    const bool plunge = !(d2x * d2x + d2y * d2y + d2z * d2z <
                          q.swallow_radius * q.swallow_radius);
    if (plunge) {
        (void)q_acc_kind(a, s, j, q, r, rc);
        q_acc_effects(a, s, j, rc, q);
    }
    d->swallowed += a.m[j];
    const double M = a.m[s] + a.m[j];
    a.vx[s] = (a.m[s] * a.vx[s] + a.m[j] * a.vx[j]) / M;
    a.vy[s] = (a.m[s] * a.vy[s] + a.m[j] * a.vy[j]) / M;
    a.vz[s] = (a.m[s] * a.vz[s] + a.m[j] * a.vz[j]) / M;
    const double rx = a.x[j] - a.x[s], ry = a.y[j] - a.y[s], rz = a.z[j] - a.z[s];
    const double ux = a.vx[j] - a.vx[s], uy = a.vy[j] - a.vy[s], uz = a.vz[j] - a.vz[s];
    const double lx = ry * uz - rz * uy, ly = rz * ux - rx * uz, lz = rx * uy - ry * ux;
    const double l = sqrt(lx * lx + ly * ly + lz * lz);
    Body &bh = a.aos[s];
    if (l > 1e-15) {
        const double w = a.m[j] / M * 0.01;
        bh.spin_x += w * lx / l;
        bh.spin_y += w * ly / l;
        bh.spin_z += w * lz / l;
        const double sm = sqrt(bh.spin_x * bh.spin_x + bh.spin_y * bh.spin_y + bh.spin_z * bh.spin_z);
        if (sm > 1e-15) {
            bh.spin_x /= sm;
            bh.spin_y /= sm;
            bh.spin_z /= sm;
        }
    }
    a.m[s] = M;
    q_kill(j, a.m, a.vx, a.vy, a.vz, a.ax, a.ay, a.az);
    d->i0 = j + 1;
    d->first = d->n;
}

// ------------------------------------------------- luminosity and feedback
__global__ __launch_bounds__(CSG_TPB) void k_q_luminosity(QArrays a, csg_quasar_config q,
                                                          double dt, const QDev *d) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= d->n) return;
    Body &b = a.aos[i];
    if (b.type != BODY_SMBH) return;
    const double mdot = d->swallowed / dt;
    b.accretion_rate = q.accretion_smoothing * b.accretion_rate + (1.0 - q.accretion_smoothing) * mdot;
    const double ledd = q.eddington_k * a.m[i];
    b.luminosity = q.eta_eff * b.accretion_rate;
    if (b.luminosity > ledd) b.luminosity = ledd;
}

// Radiation pressure from every shining black hole, in index order (quasar.c:200-230).
__global__ __launch_bounds__(CSG_TPB) void k_q_feedback(QArrays a, csg_quasar_config q,
                                                        const int *__restrict__ list,
                                                        const QDev *d) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= d->n || a.m[i] <= 0.0 || a.aos[i].type == BODY_SMBH) return;
    const int ns = min(d->nsmbh, Q_MAX_SMBH);
    const double rfb2 = (q.accretion_radius * 3.0) * (q.accretion_radius * 3.0);
    for (int k = 0; k < ns; ++k) {
        const int s = list[k];
        const double lum = a.aos[s].luminosity;
        if (lum <= 0.0) continue;
        const double dx = a.x[i] - a.x[s], dy = a.y[i] - a.y[s], dz = a.z[i] - a.z[s];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > rfb2 || d2 < 1e-10) continue;
        const double r = sqrt(d2);
        const double f = q.feedback_strength * lum / (4.0 * SPH_PI * d2);
        a.ax[i] += f * dx / r;
        a.ay[i] += f * dy / r;
        a.az[i] += f * dz / r;
    }
}

// ------------------------------------------------------------ ranks
// Bodies that draw from the random stream or take new slots are ranked by index:
// per-block counts of a flag array, a scan of them (k_gas_scan), then each
// thread's rank is the block's offset plus a block scan of its flag.
__global__ __launch_bounds__(CSG_TPB) void k_q_tiles(const int *__restrict__ flag, int n,
                                                     int *__restrict__ tilecnt) {
    __shared__ int lds4[4];
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    int total;
    (void)block_excl_scan(i < n ? flag[i] : 0, lds4, total);
    if (threadIdx.x == 0) tilecnt[blockIdx.x] = total;
}

__device__ __forceinline__ int q_rank(const int *flag, int n, const int *tilecnt, int *lds4) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    int total;
    return tilecnt[blockIdx.x] + block_excl_scan(i < n ? flag[i] : 0, lds4, total);
}

// One thread: `per` draws for each of the phase's `count` bodies (capped).
__global__ void k_q_draws(QDev *d, const int *count, int cap_bodies, int per, double *out) {
    const int c = min(*count, cap_bodies);
    d->cnt = c;
    unsigned long long st = d->rng;
    for (int t = 0; t < c * per; ++t) out[t] = q_uniform(st);
    d->rng = st;
}

// ------------------------------------------------------------ jet decay
// Live jets age; a jet at the end of its life dies, and one heavier than 0.2
// leaves a lobe behind (quasar.c:393-418).  flag[i] = dying jet with a lobe.
__global__ __launch_bounds__(CSG_TPB) void k_q_decay_flag(QArrays a, double dt, int n0,
                                                          int *__restrict__ flag,
                                                          int *__restrict__ dying) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n0) return;
    int f = 0, dd = 0;
    Body &b = a.aos[i];
    if (b.type == BODY_JET && a.m[i] > 0.0) {
        b.lifetime -= dt;
        if (b.lifetime <= 0.0) {
            dd = 1;
            f = a.m[i] > 0.2 ? 1 : 0;
        }
    }
    flag[i] = f;
    dying[i] = dd;
}

__global__ __launch_bounds__(CSG_TPB) void k_q_decay_apply(QArrays a, csg_quasar_config q, int n0,
                                                           int n_alloc, const int *__restrict__ flag,
                                                           const int *__restrict__ dying,
                                                           const int *__restrict__ tilecnt,
                                                           const double *__restrict__ u, QDev *d) {
    __shared__ int lds4[4];
    const int rank = q_rank(flag, n0, tilecnt, lds4);
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n0 || !dying[i]) return;
    if (flag[i] && n0 + rank < n_alloc) {
        const int k = n0 + rank;
        Body lobe = {};
        lobe.type = BODY_LOBE;
        lobe.lifetime = q.jet_lifetime * 3.0;
        a.aos[k] = lobe;
        a.m[k] = a.m[i] * 0.5;
        a.x[k] = a.x[i];
        a.y[k] = a.y[i];
        a.z[k] = a.z[i];
        a.vx[k] = a.vx[i] * 0.15 + (u[3 * rank] - 0.5) * 0.5;
        a.vy[k] = a.vy[i] * 0.15 + (u[3 * rank + 1] - 0.5) * 0.5;
        a.vz[k] = a.vz[i] * 0.15 + (u[3 * rank + 2] - 0.5) * 0.5;
        a.ax[k] = a.ay[k] = a.az[k] = 0.0;
    }
    q_kill(i, a.m, a.vx, a.vy, a.vz, a.ax, a.ay, a.az);
    atomicSub(&d->jet_count, 1);
}

// Lobes (new ones included) age and die (quasar.c:421-431).
__global__ __launch_bounds__(CSG_TPB) void k_q_lobes(QArrays a, double dt, const QDev *d) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= d->n) return;
    Body &b = a.aos[i];
    if (b.type != BODY_LOBE || a.m[i] <= 0.0) return;
    b.lifetime -= dt;
    if (b.lifetime <= 0.0) q_kill(i, a.m, a.vx, a.vy, a.vz, a.ax, a.ay, a.az);
}

__global__ void k_q_grow(QDev *d, int n0, int n_alloc) {
    d->n = min(n0 + d->cnt, n_alloc);
}

// ------------------------------------------------------------ recycling
// Live jets beyond ten accretion radii of the first live black hole return as
// infalling gas on the accretion sphere (quasar.c:434-475).
__global__ __launch_bounds__(CSG_TPB) void k_q_recycle_flag(QArrays a, csg_quasar_config q,
                                                            const QDev *d, int n_alloc,
                                                            int *__restrict__ flag) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n_alloc) return;   // flags up to the capacity: the rank pass covers it
    int f = 0;
    const int s = d->s_recycle;
    if (i < d->n && s >= 0 && a.aos[i].type == BODY_JET && a.m[i] > 0.0) {
        const double rr = q.accretion_radius * 10.0;
        const double dx = a.x[i] - a.x[s], dy = a.y[i] - a.y[s], dz = a.z[i] - a.z[s];
        f = (dx * dx + dy * dy + dz * dz < rr * rr) ? 0 : 1;
    }
    flag[i] = f;
}

__global__ __launch_bounds__(CSG_TPB) void k_q_recycle_apply(QArrays a, csg_quasar_config q,
                                                             const int *__restrict__ flag,
                                                             const int *__restrict__ tilecnt,
                                                             const double *__restrict__ u, QDev *d) {
    __shared__ int lds4[4];
    const int n = d->n;
    const int rank = q_rank(flag, n, tilecnt, lds4);
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const int s = d->s_recycle;
    const double angle = 2.0 * SPH_PI * u[2 * rank];
    const double cphi = 2.0 * u[2 * rank + 1] - 1.0;
    const double sphi = sqrt(1.0 - cphi * cphi);
    const double rp = q.accretion_radius;
    a.x[i] = a.x[s] + rp * sphi * cos(angle);
    a.y[i] = a.y[s] + rp * sphi * sin(angle);
    a.z[i] = a.z[s] + rp * cphi;
    const double vin = 0.5 * sqrt(a.m[s] / rp);
    const double nx = (a.x[i] - a.x[s]) / rp, ny = (a.y[i] - a.y[s]) / rp,
                 nz = (a.z[i] - a.z[s]) / rp;
    a.vx[i] = a.vx[s] - vin * nx;
    a.vy[i] = a.vy[s] - vin * ny;
    a.vz[i] = a.vz[s] - vin * nz;
    a.ax[i] = a.ay[i] = a.az[i] = 0.0;
    Body &b = a.aos[i];
    b.type = BODY_GAS;
    b.lifetime = 0.0;
    b.smoothing_h = q.accretion_radius * 0.3;
    b.internal_energy = 0.001 * (a.vx[i] * a.vx[i] + a.vy[i] * a.vy[i] + a.vz[i] * a.vz[i]) + 0.01;
    b.density = 0.0;
    b.pressure = 0.0;
    atomicSub(&d->jet_count, 1);
}

// ------------------------------------------------------------ spawning
// One thread, the reference's order (quasar.c:232-391): every shining black
// hole emits pulses (two to eight, plus four burst knots at speed peaks), each a
// ring of jet bodies, while the jet cap and the capacity allow.
__global__ void k_q_spawn(QArrays a, csg_quasar_config q, int phase, int n_alloc,
                          const int *__restrict__ list, QDev *d) {
    unsigned long long st = d->rng;
    int n = d->n, jets = d->jet_count;
    double smod = 1.0 + 0.6 * sin(phase * 0.10);
    if (smod < 0.5) smod = 0.5;
    const double cone = 0.05 + 0.13 * (0.5 + 0.5 * sin(phase * 0.057));
    const double hph = phase * 0.04;
    const int ns = min(d->nsmbh, Q_MAX_SMBH);
    for (int k = 0; k < ns; ++k) {
        const int s = list[k];
        const Body &bh = a.aos[s];
        if (bh.luminosity <= 0.0) continue;
        const double energy = 0.5 * q.jet_mass * q.jet_speed * q.jet_speed;
        if (energy < 1e-15) continue;
        int pulses = (int)(bh.luminosity / energy);
        if (pulses < 2) pulses = 2;
        if (pulses > 8) pulses = 8;
        double px, py, pz;
        if (fabs(bh.spin_z) < 0.9) {
            px = -bh.spin_y; py = bh.spin_x; pz = 0.0;
        } else {
            px = 0.0; py = -bh.spin_z; pz = bh.spin_y;
        }
        const double pm = sqrt(px * px + py * py + pz * pz);
        if (pm > 1e-15) {
            px /= pm; py /= pm; pz /= pm;
        }
        double qx = bh.spin_y * pz - bh.spin_z * py;
        double qy = bh.spin_z * px - bh.spin_x * pz;
        double qz = bh.spin_x * py - bh.spin_y * px;
        const double qm = sqrt(qx * qx + qy * qy + qz * qz);
        if (qm > 1e-15) {
            qx /= qm; qy /= qm; qz /= qm;
        }
        const double ha = q.swallow_radius * 0.35;
        const double hx = ha * (cos(hph) * px + sin(hph) * qx);
        const double hy = ha * (cos(hph) * py + sin(hph) * qy);
        const double hz = ha * (cos(hph) * pz + sin(hph) * qz);
        const double speed = q.jet_speed * smod;
        const int total = pulses + ((smod > 1.35) ? 4 : 0);
        for (int j = 0; j < total && jets < q.jet_cap; ++j) {
            if (n >= n_alloc) break;
            const bool burst = j >= pulses;
            const double sg = (j % 2 == 0) ? 1.0 : -1.0;
            const double cn = burst ? 0.012 : cone;
            const double sp = burst ? speed * 1.28 : speed;
            const double ms = burst ? q.jet_mass * 1.9 : q.jet_mass;
            const double life = burst ? q.jet_lifetime * 1.4 : q.jet_lifetime;
            const double fwd = burst ? q.swallow_radius * 0.6 : 0.0;
            const double ang = q_uniform(st) * 2.0 * SPH_PI;
            double u1 = q_uniform(st);
            const double u2 = q_uniform(st);
            if (u1 < 1e-15) u1 = 1e-15;
            const double off = sqrt(-2.0 * log(u1)) * cos(2.0 * SPH_PI * u2) * cn;
            const int ring = burst ? 1 : q.jet_ring_count;
            for (int r = 0; r < ring; ++r) {
                if (n >= n_alloc || jets >= q.jet_cap) break;
                Body e = {};
                e.type = BODY_JET;
                e.lifetime = life;
                const double ra = (ring > 1) ? 2.0 * SPH_PI * r / ring + ang : ang;
                const double rr = (ring > 1) ? q.swallow_radius * 0.15 : 0.0;
                const double ox = rr * (cos(ra) * px + sin(ra) * qx);
                const double oy = rr * (cos(ra) * py + sin(ra) * qy);
                const double oz = rr * (cos(ra) * pz + sin(ra) * qz);
                a.aos[n] = e;
                a.m[n] = (ring > 1) ? ms * 0.6 : ms;
                a.x[n] = a.x[s] + sg * bh.spin_x * (q.swallow_radius + fwd) + off * (px * cos(ang)) +
                         sg * hx + ox;
                a.y[n] = a.y[s] + sg * bh.spin_y * (q.swallow_radius + fwd) + off * (py * cos(ang)) +
                         sg * hy + oy;
                a.z[n] = a.z[s] + sg * bh.spin_z * (q.swallow_radius + fwd) + off * (pz * sin(ang)) +
                         sg * hz + oz;
                const double ev = sp * 0.04 * (1.0 + life * 0.1);
                a.vx[n] = a.vx[s] + sg * bh.spin_x * sp + ox * ev / (rr + 1e-15);
                a.vy[n] = a.vy[s] + sg * bh.spin_y * sp + oy * ev / (rr + 1e-15);
                a.vz[n] = a.vz[s] + sg * bh.spin_z * sp + oz * ev / (rr + 1e-15);
                a.ax[n] = a.ay[n] = a.az[n] = 0.0;
                ++n;
                ++jets;
            }
        }
    }
    d->rng = st;
    d->n = n;
    d->jet_count = jets;
}

// ------------------------------------------------------------ compaction
// Live bodies (mass > 0) keep their order and move down (quasar.c:512-523):
// every array is gathered into scratch and copied back.
__global__ __launch_bounds__(CSG_TPB) void k_q_live_flag(const double *__restrict__ m, int n,
                                                         int *__restrict__ flag) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i < n) flag[i] = m[i] > 0.0 ? 1 : 0;
}

__global__ __launch_bounds__(CSG_TPB) void k_q_compact(QArrays a, QArrays t, int n,
                                                       const int *__restrict__ flag,
                                                       const int *__restrict__ tilecnt) {
    __shared__ int lds4[4];
    const int k = q_rank(flag, n, tilecnt, lds4);
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n || !flag[i]) return;
    t.x[k] = a.x[i]; t.y[k] = a.y[i]; t.z[k] = a.z[i];
    t.vx[k] = a.vx[i]; t.vy[k] = a.vy[i]; t.vz[k] = a.vz[i];
    t.ax[k] = a.ax[i]; t.ay[k] = a.ay[i]; t.az[k] = a.az[i];
    t.m[k] = a.m[i];
    t.aos[k] = a.aos[i];
}

// ---------------------------------------------------------------- host side
static void q_free(csg_ctx *c) {
    if (c->qtmp) (void)hipFree(c->qtmp);
    c->qtmp = nullptr;
    c->qtmp_cap = 0;
    if (c->qbuf) (void)hipFree(c->qbuf);
    if (c->qdev) (void)hipFree(c->qdev);
    if (c->qhost) (void)hipHostFree(c->qhost);
    c->qbuf = nullptr;
    c->qdev = nullptr;
    c->qhost = nullptr;
    c->qbuf_cap = 0;
}

static QArrays q_arrays(csg_ctx *c) {
    return QArrays{c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->m, c->aos};
}

// Scratch for the context's capacity: flags, dying marks, tile counts, the
// black-hole list and three draws per body.
static int q_prepare(csg_ctx *c) {
    const size_t cap = (size_t)c->cap + 1;
    const size_t nb = cap / CSG_TPB + 2;
    const size_t need = sizeof(int) * (2 * cap + nb + 64 + Q_MAX_SMBH) + sizeof(double) * 3 * cap;
    int s = grow(&c->qbuf, &c->qbuf_cap, need);
    if (s) return s;
    if (!c->qdev && hipMalloc((void **)&c->qdev, sizeof(QDev)) != hipSuccess) {
        (void)hipGetLastError();
        c->qdev = nullptr;
        return COSMOSIM_E_ALLOC;
    }
    if (!c->qhost && hipHostMalloc((void **)&c->qhost, sizeof(QDev), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->qhost = nullptr;
        return COSMOSIM_E_ALLOC;
    }
    return COSMOSIM_OK;
}

struct QScratch {
    double *u;
    int *flag, *dying, *tilecnt, *total, *list;
};

// tilecnt holds up to nb block counts (every scan here covers at most the
// capacity); the scan's total sits after all of them.
static QScratch q_scratch(csg_ctx *c) {
    const size_t cap = (size_t)c->cap + 1;
    const size_t nb = cap / CSG_TPB + 2;
    QScratch q;
    q.u = (double *)c->qbuf;
    q.flag = (int *)(q.u + 3 * cap);
    q.dying = q.flag + cap;
    q.tilecnt = q.dying + cap;
    q.total = q.tilecnt + nb + 32;
    q.list = q.tilecnt + nb + 64;
    return q;
}

static int q_read(csg_ctx *c) {
    CSG_CHECK(hipMemcpyAsync(c->qhost, c->qdev, sizeof(QDev), hipMemcpyDeviceToHost, c->stream));
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

// Rank pass: block counts of flag[0..n) into tilecnt and their scan; the total
// goes to *total.
static void q_scan(csg_ctx *c, const int *flag, int n, int *tilecnt, int *total) {
    const int nb = at_least1(blocks_for(n, CSG_TPB));
    k_q_tiles<<<nb, CSG_TPB, 0, c->stream>>>(flag, n, tilecnt);
    k_gas_scan<<<1, CSG_TPB, 0, c->stream>>>(tilecnt, nb, total);
}

extern "C" {

// One quasar step on the device state (src/quasar.c:477-510); cfg->jet_count is
// updated and the body count may grow (jets, lobes) up to the capacity.
int csg_quasar_step(csg_ctx *c, csg_quasar_config *cfg, double dt) {
    CSG_USE(c);
    if (!cfg || !(dt > 0.0)) return COSMOSIM_E_ARG;
    int s = q_prepare(c);
    if (s) return s;
    const int n0 = c->n, n_alloc = c->cap;
    hipStream_t st = c->stream;
    QArrays a = q_arrays(c);
    QScratch w = q_scratch(c);
    const csg_quasar_config q = *cfg;
    QDev h = {};
    h.rng = c->qrng;
    h.n = n0;
    h.jet_count = cfg->jet_count;
    h.s_recycle = -1;
    *c->qhost = h;
    CSG_CHECK(hipMemcpyAsync(c->qdev, c->qhost, sizeof(QDev), hipMemcpyHostToDevice, st));
    const int nb0 = at_least1(blocks_for(n0, CSG_TPB));
    k_q_list<<<nb0, CSG_TPB, 0, st>>>(a, n0, w.list, c->qdev);
    k_q_sort<<<1, 1, 0, st>>>(a, w.list, c->qdev);
    CSG_LAUNCHED();
    if ((s = q_read(c))) return s;
    const int ns = c->qhost->nsmbh;
    if (ns > Q_MAX_SMBH) return COSMOSIM_E_CAPACITY;
    int list[Q_MAX_SMBH];
    if (ns > 0)
        CSG_CHECK(hipMemcpy(list, w.list, sizeof(int) * (size_t)ns, hipMemcpyDeviceToHost));
    // 1. accretion, black hole by black hole, in passes between events
    for (int k = 0; k < ns; ++k) {
        const int bh = list[k];
        k_acc_begin<<<1, 1, 0, st>>>(a, bh, c->qdev);
        for (;;) {
            for (int it = 0; it < 8; ++it) {
                k_acc_first<<<nb0, CSG_TPB, 0, st>>>(a, bh, q, c->qdev);
                k_acc_apply<<<nb0, CSG_TPB, 0, st>>>(a, bh, q, c->qdev);
                k_acc_event<<<1, 1, 0, st>>>(a, bh, q, c->qdev);
            }
            CSG_LAUNCHED();
            if ((s = q_read(c))) return s;
            if (c->qhost->done) break;
        }
    }
    // 2. accretion rate and luminosity, 3. feedback
    k_q_luminosity<<<nb0, CSG_TPB, 0, st>>>(a, q, dt, c->qdev);
    k_q_feedback<<<nb0, CSG_TPB, 0, st>>>(a, q, w.list, c->qdev);
    // 4. jet decay: lobes ranked by index take 3 draws each, then lobes age
    k_q_decay_flag<<<nb0, CSG_TPB, 0, st>>>(a, dt, n0, w.flag, w.dying);
    int *total = w.total;
    q_scan(c, w.flag, n0, w.tilecnt, total);
    k_q_draws<<<1, 1, 0, st>>>(c->qdev, total, n_alloc - n0, 3, w.u);
    k_q_decay_apply<<<nb0, CSG_TPB, 0, st>>>(a, q, n0, n_alloc, w.flag, w.dying, w.tilecnt, w.u,
                                             c->qdev);
    k_q_grow<<<1, 1, 0, st>>>(c->qdev, n0, n_alloc);
    const int nb1 = at_least1(blocks_for(n_alloc, CSG_TPB));
    k_q_lobes<<<nb1, CSG_TPB, 0, st>>>(a, dt, c->qdev);
    // 4b. recycling of distant jets: 2 draws each, by index
    k_q_recycle_flag<<<nb1, CSG_TPB, 0, st>>>(a, q, c->qdev, n_alloc, w.flag);
    q_scan(c, w.flag, n_alloc, w.tilecnt, total);
    k_q_draws<<<1, 1, 0, st>>>(c->qdev, total, n_alloc, 2, w.u);
    k_q_recycle_apply<<<nb1, CSG_TPB, 0, st>>>(a, q, w.flag, w.tilecnt, w.u, c->qdev);
    // 5. spawning
    k_q_spawn<<<1, 1, 0, st>>>(a, q, ++c->qphase, n_alloc, w.list, c->qdev);
    CSG_LAUNCHED();
    if ((s = q_read(c))) return s;
    c->n = c->qhost->n;
    c->qrng = c->qhost->rng;
    cfg->jet_count = c->qhost->jet_count;
    c->tree_valid = 0;
    c->counts_valid = 0;
    c->wcost_n = 0;
    return COSMOSIM_OK;
}

// quasar_compact on the device state (src/quasar.c:512-523); returns the new
// count through *n_out.
int csg_quasar_compact(csg_ctx *c, int *n_out) {
    CSG_USE(c);
    int s = q_prepare(c);
    if (s) return s;
    const int n = c->n;
    hipStream_t st = c->stream;
    QScratch w = q_scratch(c);
    // scratch arrays for the gather: the build's sorted copies and buffers
    // sized for the capacity (none holds state between substeps)
    const size_t cap = (size_t)c->cap;
    if ((s = grow(&c->qtmp, &c->qtmp_cap, (sizeof(double) * 10 + sizeof(Body)) * cap + 256)))
        return s;
    double *T = (double *)c->qtmp;
    QArrays t{T, T + cap, T + 2 * cap, T + 3 * cap, T + 4 * cap, T + 5 * cap, T + 6 * cap,
              T + 7 * cap, T + 8 * cap, T + 9 * cap, (Body *)(T + 10 * cap)};
    const int nb = at_least1(blocks_for(n, CSG_TPB));
    int *total = w.total;
    k_q_live_flag<<<nb, CSG_TPB, 0, st>>>(c->m, n, w.flag);
    q_scan(c, w.flag, n, w.tilecnt, total);
    k_q_compact<<<nb, CSG_TPB, 0, st>>>(q_arrays(c), t, n, w.flag, w.tilecnt);
    CSG_LAUNCHED();
    int live = 0;
    CSG_CHECK(hipMemcpyAsync(&live, total, sizeof(int), hipMemcpyDeviceToHost, st));
    CSG_CHECK(hipStreamSynchronize(st));
    double *dst[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->m};
    for (int k = 0; k < 10; ++k)
        CSG_CHECK(hipMemcpyAsync(dst[k], T + (size_t)k * cap, sizeof(double) * (size_t)live,
                                 hipMemcpyDeviceToDevice, st));
    CSG_CHECK(hipMemcpyAsync(c->aos, t.aos, sizeof(Body) * (size_t)live, hipMemcpyDeviceToDevice, st));
    CSG_CHECK(hipStreamSynchronize(st));
    c->n = live;
    c->tree_valid = 0;
    c->counts_valid = 0;
    c->wcost_n = 0;
    if (n_out) *n_out = live;
    return COSMOSIM_OK;
}

// The random stream and knot phase of the device quasar step (the reference's
// file-scope statics start at 1 and 0).
int csg_quasar_state(csg_ctx *c, unsigned long long *rng, int *phase, int set) {
    if (!c) return COSMOSIM_E_ARG;
    if (set) {
        if (rng) c->qrng = *rng;
        if (phase) c->qphase = *phase;
    } else {
        if (rng) *rng = c->qrng;
        if (phase) *phase = c->qphase;
    }
    return COSMOSIM_OK;
}

}  // extern "C"
