// csg_neighbors.inl - batched sphere queries over the device tree (the GPU form
// of src/octree.c:236-318, octree_find_neighbors).  Included at the end of
// csg_core.hip (same translation unit as the context and the build).
//
// Semantics, per query (centre q, radius r; SPH queries around its bodies):
//   visit nodes depth first, children in octant-slot order (the pre-order of the
//   node array); a node whose cell box [c - h, c + h] is farther than r from q
//   (ex = max(0, |q - c| - h) per axis, ex^2 + ey^2 + ez^2 > r^2) is pruned; a
//   leaf reports its body when that body is live and |x_b - q|^2 <= r^2; the
//   result holds the first `cap` bodies so reported, in visiting order.  The FP64
//   expressions and their order are the reference's (no contraction), so the
//   lists are identical, cap included (src/octree.c:251-281).
//
// Layout: one wavefront per 64 queries taken in Morton order of the query centres
// (a radix sort of the slots by their 10-level cell path), so a wave's queries are
// neighbours and share most of their traversal; the traversal is the force walk's
// stackless wave-shared one (a lane takes part in node ii while its position
// p <= ii).  The order only affects the speed, never a result.

// Geometric centre and half size of every node: the FP64 cell recurrence along
// the path of the node's first member (as the reference pool export).
__global__ __launch_bounds__(CSG_TPB) void k_node_geom(const WNode *__restrict__ nodes,
                                                       const uint64_t *__restrict__ key,
                                                       const uint64_t *__restrict__ ext,
                                                       const BuildParams *__restrict__ bp,
                                                       double4 *__restrict__ geom) {
    const int v = blockIdx.x * CSG_TPB + threadIdx.x;
    if (v >= bp->n_nodes) return;
    double cx = bp->root[0], cy = bp->root[1], cz = bp->root[2], h = bp->root[3];
    if (bp->n_live > 0) {
        const WNode w = nodes[v];
        for (int l = 0; l < w.depth; ++l) cell_child(digit_at(key, ext, w.first, l), cx, cy, cz, h);
    }
    geom[v] = make_double4(cx, cy, cz, h);
}

// Sort keys of the query slots: the first 10 levels of the centre's cell path.
__global__ __launch_bounds__(CSG_TPB) void k_nb_keys(const double *__restrict__ qc, int k,
                                                     const BuildParams *__restrict__ bp,
                                                     uint64_t *__restrict__ qkey,
                                                     int *__restrict__ qslot) {
    const int q = blockIdx.x * CSG_TPB + threadIdx.x;
    if (q >= k) return;
    double cx = bp->root[0], cy = bp->root[1], cz = bp->root[2], h = bp->root[3];
    qkey[q] = path_word(qc[3 * (size_t)q], qc[3 * (size_t)q + 1], qc[3 * (size_t)q + 2], cx, cy,
                        cz, h, 10);
    qslot[q] = q;
}

__device__ __forceinline__ double box_gap(double q, double c, double h) {
    return fmax(0.0, fabs(q - c) - h);
}

__global__ __launch_bounds__(CSG_TPB) void k_neighbors(
    const WNode32 *__restrict__ n32, const double4 *__restrict__ geom,
    const BuildParams *__restrict__ bp, const double *__restrict__ px,
    const double *__restrict__ py, const double *__restrict__ pz, const double *__restrict__ pm,
    const double *__restrict__ qc, const double *__restrict__ qrad,
    const int *__restrict__ slots, int k, int cap, int *__restrict__ out,
    int *__restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * CSG_TPB + threadIdx.x) >> 6));
    if (w * 64 >= k) return;
    const int nn = bp->overflow ? 0 : bp->n_nodes;
    const int s = w * 64 + lane;
    const bool act = s < k;
    const int q = act ? slots[s] : -1;
    const double qx = act ? qc[3 * (size_t)q] : 0.0, qy = act ? qc[3 * (size_t)q + 1] : 0.0,
                 qz = act ? qc[3 * (size_t)q + 2] : 0.0;
    const double rad = act ? qrad[q] : 0.0;
    const double r2 = rad * rad;
    int p = (act && cap > 0) ? 0 : 0x7fffffff;
    int cnt = 0;
    int *const dst = act ? out + (size_t)q * cap : nullptr;
    int ii = 0;
    while (ii < nn) {
        const double4 g = geom[ii];
        const WNode32 nd = n32[ii];
        const double ex = box_gap(qx, g.x, g.w);
        const double ey = box_gap(qy, g.y, g.w);
        const double ez = box_gap(qz, g.z, g.w);
        const bool here = p <= ii;
        const bool inside = !(ex * ex + ey * ey + ez * ez > r2);   // src/octree.c:254-256
        const bool visit = here && inside;
        unsigned long long D = 0;
        if (nd.body >= 0) {
            // leaf: test the body itself (src/octree.c:259-272); every lane here moves on
            if (visit) {
                const int bi = nd.body;
                if (!(pm[bi] <= 0.0)) {
                    const double bx = px[bi] - qx, by = py[bi] - qy, bz = pz[bi] - qz;
                    if (bx * bx + by * by + bz * bz <= r2) {
                        dst[cnt] = bi;
                        ++cnt;
                    }
                }
            }
            if (here) p = ii + 1;
            if (cnt >= cap) p = 0x7fffffff;   // full: the reference stops visiting
        } else {
            // internal: a pruned cell is skipped, an intersected one is entered
            if (here && !inside) p = nd.skip;
            D = __builtin_amdgcn_ballot_w64(visit);
        }
        const int jump = max(nd.skip, ii + 1);   // strictly increasing: the loop ends
        ii = D ? ii + 1 : jump;
        if (__builtin_amdgcn_ballot_w64(p != 0x7fffffff) == 0ull) break;   // every lane full
    }
    if (act) counts[q] = cnt;
}

// ---------------------------------------------------------------- host side
// A device buffer that only grows.
static int grow(void **p, size_t *cap, size_t bytes) {
    if (bytes <= *cap && *p) return COSMOSIM_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return COSMOSIM_E_ALLOC;
    }
    *cap = bytes;
    return COSMOSIM_OK;
}

static void nb_free(csg_ctx *c) {
    if (c->geom) (void)hipFree(c->geom);
    if (c->nbuf) (void)hipFree(c->nbuf);
    c->geom = nullptr; c->nbuf = nullptr;
    c->geom_cap = c->nbuf_cap = 0;
    c->geom_valid = 0;
}

extern "C" {

// Sphere queries with centres centres[k][3] and radii radius[k] over the current
// tree; out[q * cap ..] receives the first counts[q] (<= cap) reported bodies of
// query q (original body indices) in the reference's visiting order.
int csg_find_neighbors(csg_ctx *c, const double *centres, const double *radius, int k, int cap,
                       int *out, int *counts) {
    CSG_USE(c);
    if (k < 0 || cap < 0 || (k > 0 && (!centres || !radius || !counts || (cap > 0 && !out))))
        return COSMOSIM_E_ARG;
    if (!c->tree_valid) return COSMOSIM_E_STATE;
    if ((size_t)k * (size_t)cap > (size_t)INT32_MAX) return COSMOSIM_E_ARG;
    if (k == 0) return COSMOSIM_OK;
    int s = ensure_nodes64(c, COSMOSIM_PREC_FP64);
    if (s) return s;
    if ((s = read_params(c))) return s;
    if (c->hbp->overflow) return COSMOSIM_E_CAPACITY;
    const int nn = c->hbp->n_nodes;
    hipStream_t st = c->stream;
    if (!c->geom_valid) {
        if ((s = grow((void **)&c->geom, &c->geom_cap, sizeof(double4) * (size_t)nn))) return s;
        k_node_geom<<<blocks_for(nn, CSG_TPB), CSG_TPB, 0, st>>>(c->nodes, c->key, c->ext, c->bp,
                                                                 c->geom);
        CSG_LAUNCHED();
        c->geom_valid = 1;
    }
    // scratch: centres[k][3], radii[k], keys 2 x [k], slots 2 x [k], hist, tot, out, counts
    const int nbq = at_least1(blocks_for(k, CSG_TILE));
    const size_t a = 256;
    auto up = [&](size_t bytes) { return (bytes + a - 1) / a * a; };
    const size_t o_qc = 0, o_qrad = o_qc + up(sizeof(double) * 3 * (size_t)k),
                 o_k0 = o_qrad + up(sizeof(double) * (size_t)k),
                 o_k1 = o_k0 + up(sizeof(uint64_t) * (size_t)k),
                 o_s0 = o_k1 + up(sizeof(uint64_t) * (size_t)k),
                 o_s1 = o_s0 + up(sizeof(int) * (size_t)k),
                 o_hist = o_s1 + up(sizeof(int) * (size_t)k),
                 o_tot = o_hist + up(sizeof(int) * 256 * (size_t)nbq),
                 o_out = o_tot + up(sizeof(int) * 256),
                 o_cnt = o_out + up(sizeof(int) * (size_t)k * (size_t)cap),
                 total = o_cnt + up(sizeof(int) * (size_t)k);
    if ((s = grow(&c->nbuf, &c->nbuf_cap, total))) return s;
    char *B = (char *)c->nbuf;
    double *dq = (double *)(B + o_qc), *dr = (double *)(B + o_qrad);
    uint64_t *k0 = (uint64_t *)(B + o_k0), *k1 = (uint64_t *)(B + o_k1);
    int *s0 = (int *)(B + o_s0), *s1 = (int *)(B + o_s1);
    int *hist = (int *)(B + o_hist), *tot = (int *)(B + o_tot);
    int *dout = (int *)(B + o_out), *dcnt = (int *)(B + o_cnt);
    CSG_CHECK(hipMemcpyAsync(dq, centres, sizeof(double) * 3 * (size_t)k, hipMemcpyHostToDevice,
                             st));
    CSG_CHECK(hipMemcpyAsync(dr, radius, sizeof(double) * (size_t)k, hipMemcpyHostToDevice, st));
    k_nb_keys<<<blocks_for(k, CSG_TPB), CSG_TPB, 0, st>>>(dq, k, c->bp, k0, s0);
    // LSD sort of the slots by the 30-bit cell path (four 8-bit digits)
    uint64_t *kin = k0, *kout = k1;
    int *vin = s0, *vout = s1;
    for (int pass = 0; pass < 4; ++pass) {
        k_rs_hist<<<nbq, CSG_TPB, 0, st>>>(kin, k, 8 * pass, hist, nbq);
        k_rs_rowscan<<<256, CSG_TPB, 0, st>>>(hist, nbq, tot);
        k_rs_scatter<<<nbq, CSG_TPB, 0, st>>>(kin, vin, kout, vout, k, 8 * pass, hist, tot, nbq);
        uint64_t *tk = kin; kin = kout; kout = tk;
        int *tv = vin; vin = vout; vout = tv;
    }
    timer_start(c);
    k_neighbors<<<blocks_for(k, CSG_TPB), CSG_TPB, 0, st>>>(
        c->nodes32, c->geom, c->bp, c->x, c->y, c->z, c->m, dq, dr, vin, k, cap, dout, dcnt);
    timer_stop(c, PH_NB);
    CSG_LAUNCHED();
    if (c->timing) c->nb_queries += k;
    if (cap > 0)
        CSG_CHECK(hipMemcpyAsync(out, dout, sizeof(int) * (size_t)k * cap, hipMemcpyDeviceToHost, st));
    CSG_CHECK(hipMemcpyAsync(counts, dcnt, sizeof(int) * (size_t)k, hipMemcpyDeviceToHost, st));
    CSG_CHECK(hipStreamSynchronize(st));
    collect_timers(c);
    return COSMOSIM_OK;
}

}  // extern "C"
