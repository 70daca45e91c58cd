// csg_sph.inl - smoothed-particle hydrodynamics of the gas bodies on the device
// tree: the GPU form of src/sph.c:10-231 (sph_compute_density,
// sph_compute_forces, sph_apply_cooling).  Included at the end of csg_core.hip
// after csg_neighbors.inl (same translation unit: context, build, node geometry).
//
// Each pass over the gas bodies is one wave-shared sphere traversal of the tree
// (the traversal of k_neighbors: one wavefront per 64 gas bodies taken in tree
// order, a lane takes part in node ii while its position p <= ii), and a lane
// consumes the bodies it reports in its own visiting order, at most
// SPH_MAX_NEIGHBORS of them, exactly as the reference loops over the list that
// octree_find_neighbors returned.  Every sum is sequential per body in that
// order with the reference's expressions (FP64, IEEE sqrt and divide, no
// contraction), so densities, pressures, smoothing lengths, accelerations and
// energies are bit-identical to src/sph.c on the same inputs and tree.
//
// State: positions, velocities, accelerations and masses are the device SoA
// arrays; type, smoothing length, density, pressure and internal energy live in
// the device copy of the body records (c->aos), which every download returns.

#define SPH_MAX_NEIGHBORS 128          // src/sph.h:9
#define SPH_GAMMA (5.0 / 3.0)          // src/sph.h:7
#define SPH_PI 3.14159265358979323846

// M4 cubic spline (src/sph.c:12-25) and its radial derivative (:27-41).
__device__ __forceinline__ double sph_w(double r, double h) {
    if (h < 1e-10) return 0.0;
    const double q = r / h;
    if (q > 1.0) return 0.0;
    const double norm = 8.0 / (SPH_PI * h * h * h);
    if (q > 0.5) {
        const double t = 1.0 - q;
        return norm * 2.0 * t * t * t;
    }
    return norm * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
}

__device__ __forceinline__ double sph_dw(double r, double h) {
    if (h < 1e-10) return 0.0;
    const double q = r / h;
    if (q > 1.0 || r < 1e-15) return 0.0;
    const double norm = 8.0 / (SPH_PI * h * h * h);
    if (q > 0.5) {
        const double t = 1.0 - q;
        return norm * (-6.0 * t * t) / h;
    }
    return norm * (-12.0 * q + 18.0 * q * q) / h;
}

__device__ __forceinline__ bool sph_is_gas(const Body &b, double m) {
    return b.type == BODY_GAS && live_mass(m);
}

// Wave-shared sphere traversal for the active lanes (query q, squared radius
// r2); calls visit(j, d2) for every reported body j (live, |x_j - q|^2 = d2 <=
// r2) in the lane's visiting order and stops the lane after
// SPH_MAX_NEIGHBORS reports (src/octree.c:251-281).  Returns the lane's count.
template <class Visit>
__device__ __forceinline__ int sph_traverse(const WNode32 *__restrict__ n32,
                                            const double4 *__restrict__ geom, int nn,
                                            const double *__restrict__ px,
                                            const double *__restrict__ py,
                                            const double *__restrict__ pz,
                                            const double *__restrict__ pm, bool act, double qx,
                                            double qy, double qz, double r2, Visit &&visit) {
    int p = act ? 0 : 0x7fffffff;
    int cnt = 0;
    int ii = 0;
    while (ii < nn) {
        const double4 g = geom[ii];
        const WNode32 nd = n32[ii];
        const double ex = box_gap(qx, g.x, g.w);
        const double ey = box_gap(qy, g.y, g.w);
        const double ez = box_gap(qz, g.z, g.w);
        const bool here = p <= ii;
        const bool inside = !(ex * ex + ey * ey + ez * ez > r2);
        const bool enter = here && inside;
        unsigned long long D = 0;
        if (nd.body >= 0) {
            if (enter) {
                const int j = nd.body;
                if (live_mass(pm[j])) {
                    const double bx = px[j] - qx, by = py[j] - qy, bz = pz[j] - qz;
                    const double d2 = bx * bx + by * by + bz * bz;
                    if (d2 <= r2) {
                        visit(j, d2);
                        ++cnt;
                    }
                }
            }
            if (here) p = ii + 1;
            if (cnt >= SPH_MAX_NEIGHBORS) p = 0x7fffffff;
        } else {
            if (here && !inside) p = nd.skip;
            D = __builtin_amdgcn_ballot_w64(enter);
        }
        const int jump = max(nd.skip, ii + 1);   // strictly increasing: the loop ends
        ii = D ? ii + 1 : jump;
        if (__builtin_amdgcn_ballot_w64(p != 0x7fffffff) == 0ull) break;
    }
    return cnt;
}

// Density, pressure and the adaptive smoothing length (src/sph.c:45-95).  The
// sum uses the smoothing length the pass started with; the new one is stored
// for the force pass.  The reference's r = |x_i - x_j| equals sqrt(d2) bit for
// bit (the same squares summed in the same order).
__global__ __launch_bounds__(CSG_TPB) void k_sph_density(
    const WNode32 *__restrict__ n32, const double4 *__restrict__ geom,
    const BuildParams *__restrict__ bp, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ pz, const double *__restrict__ pm,
    Body *__restrict__ aos, const int *__restrict__ gq, const int *__restrict__ nq_dev) {
    const int nq = *nq_dev;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * CSG_TPB + threadIdx.x) >> 6));
    if (w * 64 >= nq) return;
    const int nn = bp->overflow ? 0 : bp->n_nodes;
    const int s = w * 64 + lane;
    const bool act = s < nq;
    const int i = act ? gq[s] : 0;
    const double qx = act ? px[i] : 0.0, qy = act ? py[i] : 0.0, qz = act ? pz[i] : 0.0;
    const double h = act ? aos[i].smoothing_h : 0.0;
    double rho = 0.0;
    const int k = sph_traverse(n32, geom, nn, px, py, pz, pm, act, qx, qy, qz, h * h,
                               [&](int j, double d2) { rho += pm[j] * sph_w(sqrt(d2), h); });
    if (!act) return;
    Body &b = aos[i];
    if (k < 20 && h < 50.0)
        b.smoothing_h = h * 1.2;
    else if (k > 48 && h > 0.1)
        b.smoothing_h = h * 0.8;
    if (rho < 1e-15 || isnan(rho)) rho = 1e-15;
    b.density = rho;
    double u = b.internal_energy;
    if (u < 1e-15) u = 1e-15;
    b.pressure = (SPH_GAMMA - 1.0) * rho * u;
}

// Pressure gradient plus Monaghan viscosity between gas bodies, one tenth of
// it added to the acceleration (src/sph.c:99-196).
__global__ __launch_bounds__(CSG_TPB) void k_sph_forces(
    const WNode32 *__restrict__ n32, const double4 *__restrict__ geom,
    const BuildParams *__restrict__ bp, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ pz, const double *__restrict__ pm,
    const double *__restrict__ pvx, const double *__restrict__ pvy,
    const double *__restrict__ pvz, double *__restrict__ ax, double *__restrict__ ay,
    double *__restrict__ az, const Body *__restrict__ aos, const int *__restrict__ gq,
    const int *__restrict__ nq_dev) {
    const int nq = *nq_dev;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * CSG_TPB + threadIdx.x) >> 6));
    if (w * 64 >= nq) return;
    const int nn = bp->overflow ? 0 : bp->n_nodes;
    const int s = w * 64 + lane;
    const bool act = s < nq;
    const int i = act ? gq[s] : 0;
    const double qx = act ? px[i] : 0.0, qy = act ? py[i] : 0.0, qz = act ? pz[i] : 0.0;
    const double vxi = act ? pvx[i] : 0.0, vyi = act ? pvy[i] : 0.0, vzi = act ? pvz[i] : 0.0;
    const double hi = act ? aos[i].smoothing_h : 0.0;
    const double rhoi = act ? aos[i].density : 1.0, Pi = act ? aos[i].pressure : 0.0;
    const double ci = sqrt(SPH_GAMMA * Pi / (rhoi + 1e-15));
    double fx = 0.0, fy = 0.0, fz = 0.0;
    sph_traverse(n32, geom, nn, px, py, pz, pm, act, qx, qy, qz, hi * hi, [&](int j, double d2) {
        if (j == i) return;
        const Body &bj = aos[j];
        if (bj.type != BODY_GAS) return;
        const double dx = qx - px[j], dy = qy - py[j], dz = qz - pz[j];
        const double r = sqrt(d2);
        if (r < 1e-15) return;
        const double hbar = 0.5 * (hi + bj.smoothing_h);
        const double rhoj = bj.density, Pj = bj.pressure;
        const double vr = (vxi - pvx[j]) * dx + (vyi - pvy[j]) * dy + (vzi - pvz[j]) * dz;
        double visc = 0.0;
        if (vr < 0.0) {
            const double mu = hbar * vr / (r * r + 0.01 * hbar * hbar);
            const double cj = sqrt(SPH_GAMMA * Pj / (rhoj + 1e-15));
            const double cbar = 0.5 * (ci + cj), rhobar = 0.5 * (rhoi + rhoj);
            visc = (-1.0 * cbar * mu + 2.0 * mu * mu) / (rhobar + 1e-15);
        }
        const double g = sph_dw(r, hbar);
        const double gx = g * dx / r, gy = g * dy / r, gz = g * dz / r;
        const double f = pm[j] * (Pi / (rhoi * rhoi + 1e-15) + Pj / (rhoj * rhoj + 1e-15) + visc);
        fx -= f * gx;
        fy -= f * gy;
        fz -= f * gz;
    });
    if (!act) return;
    if (isnan(fx)) fx = 0.0;
    if (isnan(fy)) fy = 0.0;
    if (isnan(fz)) fz = 0.0;
    ax[i] = ax[i] + fx * 0.1;
    ay[i] = ay[i] + fy * 0.1;
    az[i] = az[i] + fz * 0.1;
}

// Implicit cooling toward the energy floor (src/sph.c:200-231).
__global__ __launch_bounds__(CSG_TPB) void k_sph_cooling(Body *__restrict__ aos,
                                                         const double *__restrict__ pm, int n,
                                                         double dt) {
    const int i = blockIdx.x * CSG_TPB + threadIdx.x;
    if (i >= n) return;
    Body &b = aos[i];
    if (!sph_is_gas(b, pm[i])) return;
    const double u = b.internal_energy, rho = b.density;
    if (u < 1e-15 || rho < 1e-15) return;
    double unew = u / (1.0 + dt * 0.05 * rho * 0.5 / sqrt(u));
    if (unew < 0.01) unew = 0.01;
    b.internal_energy = unew;
    b.pressure = (SPH_GAMMA - 1.0) * rho * unew;
}

// ------------------------------------------------------------- gas queue
// The live gas bodies in tree (sorted) order: per-tile counts, one scan of the
// tile counts, then each tile writes its bodies at its offset.  CSG_ITEMS
// consecutive sorted positions per thread.
__global__ __launch_bounds__(CSG_TPB) void k_gas_count(const int *__restrict__ idx,
                                                       const Body *__restrict__ aos,
                                                       const double *__restrict__ pm,
                                                       const BuildParams *__restrict__ bp,
                                                       int *__restrict__ tilecnt) {
    __shared__ int lds4[4];
    const int nl = bp->n_live;
    const int e0 = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int c = 0;
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = e0 + r;
        if (b < nl) {
            const int i = idx[b];
            c += sph_is_gas(aos[i], pm[i]) ? 1 : 0;
        }
    }
    int total;
    (void)block_excl_scan(c, lds4, total);
    if (threadIdx.x == 0) tilecnt[blockIdx.x] = total;
}

// One block: exclusive scan of the tile counts in place; *nq = the total.
__global__ __launch_bounds__(CSG_TPB) void k_gas_scan(int *__restrict__ tilecnt, int nt,
                                                      int *__restrict__ nq) {
    __shared__ int lds4[4];
    int carry = 0;
    for (int c0 = 0; c0 < nt; c0 += CSG_TPB) {
        const int t = c0 + threadIdx.x;
        const int v = t < nt ? tilecnt[t] : 0;
        int total;
        const int ex = block_excl_scan(v, lds4, total);
        if (t < nt) tilecnt[t] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *nq = carry;
}

__global__ __launch_bounds__(CSG_TPB) void k_gas_fill(const int *__restrict__ idx,
                                                      const Body *__restrict__ aos,
                                                      const double *__restrict__ pm,
                                                      const BuildParams *__restrict__ bp,
                                                      const int *__restrict__ tilecnt,
                                                      int *__restrict__ gq) {
    __shared__ int lds4[4];
    const int nl = bp->n_live;
    const int e0 = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int body[CSG_ITEMS];
    int c = 0;
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int b = e0 + r;
        body[r] = -1;
        if (b < nl) {
            const int i = idx[b];
            if (sph_is_gas(aos[i], pm[i])) body[r] = i;
        }
        c += body[r] >= 0 ? 1 : 0;
    }
    int total;
    int o = tilecnt[blockIdx.x] + block_excl_scan(c, lds4, total);
    for (int r = 0; r < CSG_ITEMS; ++r)
        if (body[r] >= 0) gq[o++] = body[r];
}

// ---------------------------------------------------------------- host side
static void sph_free(csg_ctx *c) {
    if (c->gq) (void)hipFree(c->gq);
    c->gq = nullptr;
    c->gq_cap = 0;
    c->gq_valid = 0;
}

// Node geometry and the gas queue for the current tree, on the stream without
// host synchronisation (buffers sized for the capacity).
static int sph_prepare(csg_ctx *c) {
    if (!c->tree_valid) return COSMOSIM_E_STATE;
    int s = ensure_nodes64(c, COSMOSIM_PREC_FP64);   // the geometry reads the FP64 records
    if (s) return s;
    hipStream_t st = c->stream;
    if (!c->geom_valid) {
        if ((s = grow((void **)&c->geom, &c->geom_cap, sizeof(double4) * (size_t)c->node_cap)))
            return s;
        k_node_geom<<<blocks_for(c->node_cap, CSG_TPB), CSG_TPB, 0, st>>>(c->nodes, c->key, c->ext,
                                                                        c->bp, c->geom);
        CSG_LAUNCHED();
        c->geom_valid = 1;
    }
    if (!c->gq_valid) {
        const int nt = at_least1(blocks_for(c->n, CSG_TILE));
        const size_t need = sizeof(int) * ((size_t)c->cap + (size_t)nt + 64);
        if ((s = grow((void **)&c->gq, &c->gq_cap, need))) return s;
        int *q = c->gq, *tilecnt = c->gq + c->cap, *nq = tilecnt + nt;
        k_gas_count<<<nt, CSG_TPB, 0, st>>>(c->idx, c->aos, c->m, c->bp, tilecnt);
        k_gas_scan<<<1, CSG_TPB, 0, st>>>(tilecnt, nt, nq);
        k_gas_fill<<<nt, CSG_TPB, 0, st>>>(c->idx, c->aos, c->m, c->bp, tilecnt, q);
        CSG_LAUNCHED();
        c->gq_valid = 1;
    }
    return COSMOSIM_OK;
}

static int launch_sph_density(csg_ctx *c) {
    int s = sph_prepare(c);
    if (s) return s;
    const int nt = at_least1(blocks_for(c->n, CSG_TILE));
    const int *nq = c->gq + c->cap + nt;
    k_sph_density<<<at_least1(blocks_for(c->n, CSG_TPB)), CSG_TPB, 0, c->stream>>>(
        c->nodes32, c->geom, c->bp, c->x, c->y, c->z, c->m, c->aos, c->gq, nq);
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

static int launch_sph_forces(csg_ctx *c) {
    int s = sph_prepare(c);
    if (s) return s;
    const int nt = at_least1(blocks_for(c->n, CSG_TILE));
    const int *nq = c->gq + c->cap + nt;
    k_sph_forces<<<at_least1(blocks_for(c->n, CSG_TPB)), CSG_TPB, 0, c->stream>>>(
        c->nodes32, c->geom, c->bp, c->x, c->y, c->z, c->m, c->vx, c->vy, c->vz, c->ax, c->ay,
        c->az, c->aos, c->gq, nq);
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

static int launch_sph_cooling(csg_ctx *c, double dt) {
    if (c->n == 0) return COSMOSIM_OK;
    k_sph_cooling<<<blocks_for(c->n, CSG_TPB), CSG_TPB, 0, c->stream>>>(c->aos, c->m, c->n, dt);
    CSG_LAUNCHED();
    return COSMOSIM_OK;
}

extern "C" {

// sph_compute_density / sph_compute_forces / sph_apply_cooling on the device
// state and the current tree (src/sph.h:11-20).
int csg_sph_density(csg_ctx *c) {
    CSG_USE(c);
    timer_start(c);
    int s = launch_sph_density(c);
    if (s) return s;
    timer_stop(c, PH_SPH);
    return check_overflow(c);
}

int csg_sph_forces(csg_ctx *c) {
    CSG_USE(c);
    timer_start(c);
    int s = launch_sph_forces(c);
    if (s) return s;
    timer_stop(c, PH_SPH);
    return check_overflow(c);
}

int csg_sph_cooling(csg_ctx *c, double dt) {
    CSG_USE(c);
    int s = launch_sph_cooling(c, dt);
    if (s) return s;
    CSG_CHECK(hipStreamSynchronize(c->stream));
    return COSMOSIM_OK;
}

// k device-resident substeps of integrator_step with SPH (src/integrator.c:12-57):
// kick + drift, build, gravity walk, SPH density and forces, kick, cooling.
// Bodies that cool are the ones the next substep's kick moves, so the kicks
// stay separate kernels here (same arithmetic as csg_substeps).
int csg_substeps_sph(csg_ctx *c, int k, double dt, double G, double softening, double theta,
                     int precision) {
    CSG_USE(c);
    c->theta_hint = theta;
    if (!prec_valid(precision) || k < 0) return COSMOSIM_E_ARG;
    if (k == 0) return COSMOSIM_OK;
    int s;
    timer_start(c);
    if ((s = launch_kick_drift(c, dt, 1))) return s;
    timer_stop(c, PH_KDK);
    for (int sub = 0; sub < k; ++sub) {
        timer_start(c);
        if ((s = launch_build(c, true, 0, hilbert_order_for(c, precision)))) return s;
        timer_stop(c, PH_BUILD);
        timer_start(c);
        if ((s = launch_walk(c, precision, G, softening, theta, TapeOut{}, false, sub + 1 == k)))
            return s;
        timer_stop(c, PH_WALK);
        timer_start(c);
        if ((s = launch_sph_density(c)) || (s = launch_sph_forces(c))) return s;
        timer_stop(c, PH_SPH);
        timer_start(c);
        if ((s = launch_kick(c, dt)) || (s = launch_sph_cooling(c, dt))) return s;
        if (sub + 1 < k && (s = launch_kick_drift(c, dt, 1))) return s;
        timer_stop(c, PH_KDK);
    }
    return check_overflow(c);
}

}  // extern "C"
