// csg_snapshot.inl - host snapshot formats produced on the device.  Included at
// the end of csg_core.hip.
//
// Both formats hold the live bodies (mass > 0) in index order, converted to
// float (round to nearest), as the consumers' own loops build them:
//   * the Unreal plugin's FCosmosimBodyGPU, 16 floats per body, at most
//     max_bodies of them (ue/.../CosmosimSubsystem.h:15-27, .cpp:12-43);
//   * the renderer's upload record, 10 floats per body: x, y, z, mass, vx, vy,
//     vz, type, internal energy, density (src/renderer.c:348-364).
// A live body's slot is its rank among the live bodies: per-tile counts, one
// scan of the tile counts, then each tile writes its records.  The caller gets
// the records with one device-to-host copy.

enum { SNAP_UE = 0, SNAP_RENDER = 1 };

__global__ __launch_bounds__(CSG_TPB) void k_live_count(const double *__restrict__ pm, int n,
                                                        int *__restrict__ tilecnt) {
    __shared__ int lds4[4];
    const int e0 = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int c = 0;
    for (int r = 0; r < CSG_ITEMS; ++r) c += (e0 + r < n && live_mass(pm[e0 + r])) ? 1 : 0;
    int total;
    (void)block_excl_scan(c, lds4, total);
    if (threadIdx.x == 0) tilecnt[blockIdx.x] = total;
}

template <int FMT>
__global__ __launch_bounds__(CSG_TPB) void k_snapshot(
    const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
    const double *__restrict__ pvx, const double *__restrict__ pvy,
    const double *__restrict__ pvz, const double *__restrict__ pm,
    const Body *__restrict__ aos, int n, const int *__restrict__ tilecnt, int max_out,
    float *__restrict__ out) {
    __shared__ int lds4[4];
    const int e0 = blockIdx.x * CSG_TILE + threadIdx.x * CSG_ITEMS;
    int c = 0;
    for (int r = 0; r < CSG_ITEMS; ++r) c += (e0 + r < n && live_mass(pm[e0 + r])) ? 1 : 0;
    int total;
    int slot = tilecnt[blockIdx.x] + block_excl_scan(c, lds4, total);
    for (int r = 0; r < CSG_ITEMS; ++r) {
        const int i = e0 + r;
        if (i >= n || !live_mass(pm[i])) continue;
        if (slot < max_out) {
            const Body &b = aos[i];
            if (FMT == SNAP_UE) {
                float *o = out + (size_t)slot * 16;
                o[0] = (float)px[i]; o[1] = (float)py[i]; o[2] = (float)pz[i];
                o[3] = (float)pm[i];
                o[4] = (float)pvx[i]; o[5] = (float)pvy[i]; o[6] = (float)pvz[i];
                o[7] = (float)b.type;
                o[8] = (float)b.internal_energy;
                o[9] = (float)b.density;
                o[10] = (float)b.lifetime;
                o[11] = (float)b.luminosity;
                o[12] = (float)b.spin_x; o[13] = (float)b.spin_y; o[14] = (float)b.spin_z;
                o[15] = 0.0f;
            } else {
                float *o = out + (size_t)slot * 10;
                o[0] = (float)px[i]; o[1] = (float)py[i]; o[2] = (float)pz[i];
                o[3] = (float)pm[i];
                o[4] = (float)pvx[i]; o[5] = (float)pvy[i]; o[6] = (float)pvz[i];
                o[7] = (float)b.type;
                o[8] = (float)b.internal_energy;
                o[9] = (float)b.density;
            }
        }
        ++slot;
    }
}

static void snap_free(csg_ctx *c) {
    if (c->snap) (void)hipFree(c->snap);
    c->snap = nullptr;
    c->snap_cap = 0;
}

extern "C" {

// Snapshot of the device state in format fmt (SNAP_UE: 16 floats per body, at
// most max_bodies bodies; SNAP_RENDER: 10 floats per body, all live bodies) into
// the host array out; *count = records written.
int csg_snapshot(csg_ctx *c, int fmt, float *out, int max_bodies, int *count) {
    CSG_USE(c);
    if (!count || (fmt != SNAP_UE && fmt != SNAP_RENDER) || max_bodies < 0) return COSMOSIM_E_ARG;
    const int n = c->n;
    *count = 0;
    if (n == 0) return COSMOSIM_OK;
    const int width = fmt == SNAP_UE ? 16 : 10;
    const int max_out = min(max_bodies, n);
    if (max_out > 0 && !out) return COSMOSIM_E_ARG;
    const int nt = blocks_for(n, CSG_TILE);
    const size_t rec_bytes = sizeof(float) * (size_t)width * (size_t)max_out;
    const size_t need = ((rec_bytes + 255) / 256) * 256 + sizeof(int) * ((size_t)nt + 64);
    int s = grow(&c->snap, &c->snap_cap, need);
    if (s) return s;
    float *recs = (float *)c->snap;
    int *tilecnt = (int *)((char *)c->snap + ((rec_bytes + 255) / 256) * 256);
    int *total = tilecnt + nt;
    hipStream_t st = c->stream;
    k_live_count<<<nt, CSG_TPB, 0, st>>>(c->m, n, tilecnt);
    k_gas_scan<<<1, CSG_TPB, 0, st>>>(tilecnt, nt, total);
    if (fmt == SNAP_UE)
        k_snapshot<SNAP_UE><<<nt, CSG_TPB, 0, st>>>(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m,
                                                   c->aos, n, tilecnt, max_out, recs);
    else
        k_snapshot<SNAP_RENDER><<<nt, CSG_TPB, 0, st>>>(c->x, c->y, c->z, c->vx, c->vy, c->vz,
                                                       c->m, c->aos, n, tilecnt, max_out, recs);
    CSG_LAUNCHED();
    int live = 0;
    CSG_CHECK(hipMemcpyAsync(&live, total, sizeof(int), hipMemcpyDeviceToHost, st));
    CSG_CHECK(hipStreamSynchronize(st));
    const int k = min(live, max_out);
    if (k > 0)
        CSG_CHECK(hipMemcpyAsync(out, recs, sizeof(float) * (size_t)width * (size_t)k,
                                 hipMemcpyDeviceToHost, st));
    CSG_CHECK(hipStreamSynchronize(st));
    *count = k;
    return COSMOSIM_OK;
}

}  // extern "C"
