// Fused per-iteration step on a tiled gather plan (schedule described in plan.h):
// assembly of lap/ela values in the COLORING build's per-entry summation order, plus
// the block-Jacobi prec (prec_init + prec_inversion) -- one launch per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "minifem.h"
#include "plan.h"

struct mf_plan {
    int op = 0, nE = 0, N = 0, nnz = 0, dim = 1;
    int ntiles = 0, block = 256, mask_words = 1;
    int lds_fp64 = 0, lds_mixed = 0;
    int n_isolated = 0;
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mf::TileHdr *tiles = nullptr;
    int32_t *tile_elems = nullptr;
    uint32_t *tile_verts = nullptr;
    int32_t *tile_nodes = nullptr;
    int4 *tile_rowinfo = nullptr;
    uint16_t *inc = nullptr;
    uint32_t *ent_mask = nullptr;
    uint16_t *ent_info = nullptr;
    int32_t *cb = nullptr;
    int32_t *isolated = nullptr;
};

namespace {

constexpr int kXcd = 8;  // MI355X: workgroups are dealt to the 8 XCDs round-robin
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

template <int MW>
struct MaskT;
template <>
struct MaskT<1> {
    using type = uint32_t;
};
template <>
struct MaskT<2> {
    using type = uint64_t;
};

template <typename M>
__device__ __forceinline__ int lowest_bit(M m)
{
    if constexpr (sizeof(M) == 8) return __builtin_ctzll(m);
    else return __builtin_ctz(m);
}

// One tile element's gradient of vertex v from its stage slot: {x, y, z}.
template <typename T, typename T2>
__device__ __forceinline__ void stage_grad(const T *sl, int v, T (&g)[3])
{
    const T2 xy = *reinterpret_cast<const T2 *>(sl + 2 * v);
    g[0] = xy.x;
    g[1] = xy.y;
    g[2] = sl[8 + v];
}

// One tile is one workgroup. blockIdx -> tile is XCD-aware: XCD x (= blockIdx % 8) walks
// the x-th contiguous eighth of the tiles in partition order, so tiles that share
// elements and nodes run on the same XCD close in time and share its L2.
//
// Global loads come in two dependent rounds: (1) everything static the tile needs --
// node and element ids, vertex lists, row info, incidence lists and this wave's entry
// masks; (2) the coordinate and coefficient gathers. Phase B then reads only LDS and
// registers, so its stores are never followed by a load (gfx950's vmcnt retires loads
// and stores in order).
template <int OP, int PREC, int BT, int MW>
__global__ void __launch_bounds__(BT)
    k_tiled(const mf::TileHdr *__restrict__ tiles, int ntiles,
            const int32_t *__restrict__ tile_elems, const uint32_t *__restrict__ tile_verts,
            const int32_t *__restrict__ tile_nodes, const int4 *__restrict__ rowinfo,
            const uint4 *__restrict__ inc8, const uint32_t *__restrict__ ent_mask,
            const uint16_t *__restrict__ ent_info, const double *__restrict__ coord,
            const double *__restrict__ coef, double *__restrict__ values,
            double *__restrict__ prec, int *__restrict__ singular)
{
    using T = typename std::conditional<PREC == MF_PREC_FP64, double, float>::type;
    using T2 = typename Vec2<T>::type;
    using M = typename MaskT<MW>::type;
    constexpr bool ELA = OP == MF_OP_ELA;
    constexpr int S = ELA ? mf::kStageEla : mf::kStageLap;
    constexpr int NA = ELA ? 9 : 1;
    constexpr int NW = BT / 64;
    constexpr int EPT = mf::kMaxTileElems / BT;                // tile elements per thread
    constexpr int NPT = (mf::kMaxTileNodes + BT - 1) / BT;     // tile nodes per thread
    constexpr int IPT = (mf::kMaxTileInc / 8 + BT - 1) / BT;   // 16-byte incidence loads
    constexpr int PB = mf::kMaxTileBlocks / NW;                // work blocks per wave
    extern __shared__ __align__(16) unsigned char lds_raw[];

    const int per = (ntiles + kXcd - 1) / kXcd;
    const int xcd = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    const int tile = xcd * per + slot;
    if (slot >= per || tile >= ntiles) return;
    const mf::TileHdr h = tiles[tile];
    const int NN = h.n_node, TE = h.n_elem, R = h.n_rows;
    const mf::TileLds L = mf::tile_lds_layout(NN, TE, R, h.n_inc, S * (int)sizeof(T));
    T *stage = reinterpret_cast<T *>(lds_raw + L.stage);
    int4 *lri = reinterpret_cast<int4 *>(lds_raw + L.rinfo);
    double *lc = reinterpret_cast<double *>(lds_raw + L.coords);
    int *lrnode = reinterpret_cast<int *>(lds_raw + L.rnode);
    uint32_t *lverts = reinterpret_cast<uint32_t *>(lds_raw + L.verts);
    uint16_t *linc = reinterpret_cast<uint16_t *>(lds_raw + L.inc);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Work blocks: diagonal block w (rows 64w..) on wave w; entry blocks m (CSR entries
    // 64m..) on the other waves while they hold at most 4 each (a diagonal block costs
    // about as much as four entry blocks), else on all waves round-robin.
    const int nDB = mf::diag_blocks(ELA, R);
    const int nEB = (h.n_ent + 63) / 64;
    const bool spread = nEB > 4 * (NW - nDB);
    const int m0 = spread ? wave : wave - nDB, mstep = spread ? NW : NW - nDB;

    // ---- round 1: static tile data
    int nodeId[NPT];
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int i = tid + r * BT;
        nodeId[r] = i < NN ? tile_nodes[h.node_start + i] : 0;
    }
    int elemId[EPT];
    uint32_t vert[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = tid + r * BT;
        elemId[r] = t < TE ? tile_elems[h.elem_start + t] : 0;
        vert[r] = t < TE ? tile_verts[h.elem_start + t] : 0u;
    }
    int4 rinf[NPT];
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int i = tid + r * BT;
        rinf[r] = i < R ? rowinfo[h.rinfo_start + i] : make_int4(0, 0, 0, 0);
    }
    uint4 incv[IPT];
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = tid + r * BT;
        incv[r] = i < h.n_inc / 8 ? inc8[h.inc_start / 8 + i] : make_uint4(0, 0, 0, 0);
    }
    uint32_t pinf[PB];
    M pm[PB];
#pragma unroll
    for (int k = 0; k < PB; k++) {
        const int idx = (m0 + k * mstep) * 64 + lane;
        pinf[k] = kNoEntry;
        pm[k] = 0;
        if (m0 >= 0 && idx < h.n_ent) {
            const int64_t e = (int64_t)h.ent_start + idx;
            pinf[k] = ent_info[e];
            if constexpr (MW == 1) {
                pm[k] = ent_mask[e];
            } else {
                const uint2 w = reinterpret_cast<const uint2 *>(ent_mask)[e];
                pm[k] = (uint64_t)w.x | (uint64_t)w.y << 32;
            }
        }
    }

    // ---- round 2: coordinate and coefficient gathers
    double cx[NPT][3];
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int i = tid + r * BT;
        cx[r][0] = cx[r][1] = cx[r][2] = 0.0;
        if (i < NN) {
            const double *p = coord + (int64_t)nodeId[r] * 3;
            cx[r][0] = p[0];
            cx[r][1] = p[1];
            cx[r][2] = p[2];
        }
    }
    double kc0[EPT], kc1[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = tid + r * BT;
        kc0[r] = ELA ? 0.25 : 1.0;  // reference constants: lambda 0.25, mu 1 | k 1
        kc1[r] = 1.0;
        if (t < TE && coef) {
            if constexpr (ELA) {
                const double2 lm = reinterpret_cast<const double2 *>(coef)[elemId[r]];
                kc0[r] = lm.x;
                kc1[r] = lm.y;
            } else {
                kc0[r] = coef[elemId[r]];
            }
        }
    }

    // ---- static data into LDS
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int i = tid + r * BT;
        if (i < R) {
            lri[i] = rinf[r];
            lrnode[i] = nodeId[r];
        }
    }
#pragma unroll
    for (int r = 0; r < IPT; r++) {
        const int i = tid + r * BT;
        if (i < h.n_inc / 8) reinterpret_cast<uint4 *>(linc)[i] = incv[r];
    }
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = tid + r * BT;
        if (t < TE) lverts[t] = vert[r];
    }
#pragma unroll
    for (int r = 0; r < NPT; r++) {
        const int i = tid + r * BT;
        if (i < NN) {
            lc[i * 3 + 0] = cx[r][0];
            lc[i * 3 + 1] = cx[r][1];
            lc[i * 3 + 2] = cx[r][2];
        }
    }
    __syncthreads();

    // ---- A: scaled gradients + coefficients of every tile element, one slot each
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = tid + r * BT;
        if (t >= TE) continue;
        const int vi[4] = {(int)(vert[r] & 0xFF), (int)((vert[r] >> 8) & 0xFF),
                           (int)((vert[r] >> 16) & 0xFF), (int)(vert[r] >> 24)};
        double c[4][3], g[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int d = 0; d < 3; d++) c[i][d] = lc[vi[i] * 3 + d];
        mf::elem_grad(c, g);
        T *sl = stage + t * S;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            *reinterpret_cast<T2 *>(sl + 2 * i) = T2{(T)g[i][0], (T)g[i][1]};
            sl[8 + i] = (T)g[i][2];
        }
        if constexpr (ELA) {
            // src/assembly.cc:386-409 with per-element (lambda, mu)
            const double cd = kc0[r] + 2.0 * kc1[r], co = kc0[r] + kc1[r], cm = kc1[r];
            *reinterpret_cast<T2 *>(sl + 12) = T2{(T)cd, (T)co};
            *reinterpret_cast<T2 *>(sl + 14) = T2{(T)cm, (T)0};
        } else {
            *reinterpret_cast<T2 *>(sl + 12) = T2{(T)kc0[r], (T)0};
        }
    }
    __syncthreads();

    // ---- B1: diagonal blocks, one lane per row: the whole incidence list, ascending,
    // four entries per LDS read with their stage reads issued together
    if (wave < nDB) {
        const int r = wave * 64 + lane;
        if (r < R) {
            const int4 ri = lri[r];
            const int incOff = ri.z & 0xFFFF, val = (ri.z >> 16) & 0xFFFF;
            const int64_t dpos = (int64_t)ri.x + ((ri.y >> 16) & 0xFF);
            if constexpr (ELA) {
                // ela_block(g, g): exactly symmetric (g_d*g_e == g_e*g_d in IEEE), so six
                // sums hold the block bit for bit
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s4 = 0.0, s5 = 0.0, s8 = 0.0;
                for (int i = 0; i < val; i += 4) {
                    const uint2 e4 = *reinterpret_cast<const uint2 *>(linc + incOff + i);
                    const int en[4] = {(int)(e4.x & 0xFFFF), (int)(e4.x >> 16),
                                       (int)(e4.y & 0xFFFF), (int)(e4.y >> 16)};
                    T g[4][3], cd[4], co[4], cm[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const T *sl = stage + (en[u] >> 2) * S;
                        stage_grad<T, T2>(sl, en[u] & 3, g[u]);
                        const T2 c01 = *reinterpret_cast<const T2 *>(sl + 12);
                        cd[u] = c01.x;
                        co[u] = c01.y;
                        cm[u] = sl[14];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (i + u >= val) break;
                        const T p00 = g[u][0] * g[u][0], p11 = g[u][1] * g[u][1],
                                p22 = g[u][2] * g[u][2];
                        s0 += (double)(p00 * cd[u] + p11 * cm[u] + p22 * cm[u]);
                        s1 += (double)(g[u][0] * g[u][1] * co[u]);
                        s2 += (double)(g[u][0] * g[u][2] * co[u]);
                        s4 += (double)(p00 * cm[u] + p11 * cd[u] + p22 * cm[u]);
                        s5 += (double)(g[u][1] * g[u][2] * co[u]);
                        s8 += (double)(p00 * cm[u] + p11 * cm[u] + p22 * cd[u]);
                    }
                }
                double blk[9] = {s0, s1, s2, s1, s4, s5, s2, s5, s8};
                double *dv = values + dpos * 9;
#pragma unroll
                for (int q = 0; q < 9; q++) dv[q] = blk[q];
                const int info = mf::ela_invert_block(blk, (ri.y >> 24) & 7, true);
                double *dp = prec + (int64_t)lrnode[r] * 9;
#pragma unroll
                for (int q = 0; q < 9; q++) dp[q] = blk[q];
                if (info != 0 && singular) atomicAdd(singular, 1);
            } else {
                double acc = 0.0;
                for (int i = 0; i < val; i += 4) {
                    const uint2 e4 = *reinterpret_cast<const uint2 *>(linc + incOff + i);
                    const int en[4] = {(int)(e4.x & 0xFFFF), (int)(e4.x >> 16),
                                       (int)(e4.y & 0xFFFF), (int)(e4.y >> 16)};
                    T g[4][3], kk[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const T *sl = stage + (en[u] >> 2) * S;
                        stage_grad<T, T2>(sl, en[u] & 3, g[u]);
                        kk[u] = sl[12];
                    }
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        if (i + u >= val) break;
                        const T x = kk[u] * mf::lap_entry(g[u], g[u]);
                        acc += (double)x;
                    }
                }
                values[dpos] = acc;
                prec[lrnode[r]] = 1.0 / acc;
            }
        }
    }

    // ---- B2: off-diagonal entries, one lane per CSR entry: the elements of the row's
    // list that hold the column, two per iteration (independent LDS chains, summed in
    // order)
#pragma clang loop unroll(disable)
    for (int k = 0; k < PB; k++) {
        const int m = m0 + k * mstep;
        if (m0 < 0 || m >= nEB) break;
        const uint32_t inf = pinf[0];
        M msk = pm[0];
#pragma unroll
        for (int s = 0; s + 1 < PB; s++) {
            pinf[s] = pinf[s + 1];
            pm[s] = pm[s + 1];
        }
        const int ra = inf & 0xFF, cl = (inf >> 8) & 0xFF;
        if (inf == kNoEntry || cl == ra) continue;  // past the end, or the diagonal (B1)
        const int4 ri = lri[ra];
        const int incOff = ri.z & 0xFFFF;
        const int oa = m * 64 + lane - ri.w;
        double acc[NA];
#pragma unroll
        for (int q = 0; q < NA; q++) acc[q] = 0.0;
        while (msk) {
            const int b0 = lowest_bit(msk);
            msk &= msk - 1;
            const bool two = msk != 0;
            const int b1 = two ? lowest_bit(msk) : b0;
            if (two) msk &= msk - 1;
            const int e0 = linc[incOff + b0], e1 = linc[incOff + b1];
            const int t[2] = {e0 >> 2, e1 >> 2}, a[2] = {e0 & 3, e1 & 3};
            T ga[2][3], gk[2][3], c0[2], c1[2], c2[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const uint32_t vv = lverts[t[u]];
                const int kk = (int)(vv & 0xFF) == cl ? 0
                               : (int)((vv >> 8) & 0xFF) == cl ? 1
                               : (int)((vv >> 16) & 0xFF) == cl ? 2 : 3;
                const T *sl = stage + t[u] * S;
                stage_grad<T, T2>(sl, a[u], ga[u]);
                stage_grad<T, T2>(sl, kk, gk[u]);
                if constexpr (ELA) {
                    const T2 c01 = *reinterpret_cast<const T2 *>(sl + 12);
                    c0[u] = c01.x;
                    c1[u] = c01.y;
                    c2[u] = sl[14];
                } else {
                    c0[u] = sl[12];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (u == 1 && !two) break;
                if constexpr (ELA) {
                    T bl[9];
                    mf::ela_block(ga[u], gk[u], c0[u], c1[u], c2[u], bl);
#pragma unroll
                    for (int q = 0; q < 9; q++) acc[q] += (double)bl[q];
                } else {
                    const T x = c0[u] * mf::lap_entry(ga[u], gk[u]);
                    acc[0] += (double)x;
                }
            }
        }
        double *dst = values + ((int64_t)ri.x + oa) * NA;
#pragma unroll
        for (int q = 0; q < NA; q++) dst[q] = acc[q];
    }
}

// Nodes with an empty row: prec_init leaves zeros; lap inverts them to inf, ela masks
// them and skips the inversion (no diagonal entry), src/Fortran/elasclpr.f:29-32.
template <int OP>
__global__ void k_prec_isolated(const int32_t *__restrict__ nodes, int n,
                                const int32_t *__restrict__ cb, int N, double *__restrict__ prec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = nodes[i];
    if (OP == MF_OP_LAP) {
        prec[node] = 1.0 / 0.0;
        return;
    }
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int mask = 0;
    if (cb)
        mask = (cb[node] != 0 ? 1 : 0) | (cb[(int64_t)N + node] != 0 ? 2 : 0) |
               (cb[2 * (int64_t)N + node] != 0 ? 4 : 0);
    mf::ela_invert_block(blk, mask, false);
    for (int q = 0; q < 9; q++) prec[(int64_t)node * 9 + q] = blk[q];
}

template <typename T>
int upload(T **dst, const std::vector<T> &src)
{
    *dst = nullptr;
    const size_t bytes = std::max<size_t>(16, src.size() * sizeof(T));
    if (hipMalloc(reinterpret_cast<void **>(dst), bytes) != hipSuccess) return MF_ERR_NOMEM;
    if (!src.empty() &&
        hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return MF_ERR_HIP;
    return MF_OK;
}

template <typename D, typename S>
int upload_as(D **dst, const std::vector<S> &src)
{
    static_assert(sizeof(D) % sizeof(S) == 0 || sizeof(S) % sizeof(D) == 0, "layout");
    std::vector<D> tmp((src.size() * sizeof(S) + sizeof(D) - 1) / sizeof(D));
    if (!src.empty()) std::memcpy(tmp.data(), src.data(), src.size() * sizeof(S));
    return upload(dst, tmp);
}

void free_plan(mf_plan *p)
{
    if (!p) return;
    void *ptrs[] = {p->tiles, p->tile_elems, p->tile_verts, p->tile_nodes, p->tile_rowinfo,
                    p->inc, p->ent_mask, p->ent_info, p->cb, p->isolated};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

template <int OP, int PREC, int BT, int MW>
int launch_tiled(mf_plan *p, const double *coord, const double *coef, double *values,
                 double *prec, int *singular, hipStream_t s)
{
    const int lds = PREC == MF_PREC_FP64 ? p->lds_fp64 : p->lds_mixed;
    auto kern = k_tiled<OP, PREC, BT, MW>;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
        return MF_ERR_HIP;
    const int per = (p->ntiles + kXcd - 1) / kXcd;
    if (p->ntiles > 0) {
        hipLaunchKernelGGL(kern, dim3(per * kXcd), dim3(BT), lds, s, p->tiles, p->ntiles,
                           p->tile_elems, p->tile_verts, p->tile_nodes, p->tile_rowinfo,
                           reinterpret_cast<const uint4 *>(p->inc), p->ent_mask, p->ent_info,
                           coord, coef, values, prec, singular);
        MF_CHECK_LAUNCH();
    }
    if (p->n_isolated > 0) {
        hipLaunchKernelGGL(k_prec_isolated<OP>, dim3((p->n_isolated + 255) / 256), dim3(256), 0, s,
                           p->isolated, p->n_isolated, p->cb, p->N, prec);
        MF_CHECK_LAUNCH();
    }
    return MF_OK;
}

template <int OP, int PREC>
int launch_variant(mf_plan *p, const double *coord, const double *coef, double *values,
                   double *prec, int *singular, hipStream_t s)
{
    const bool wide = p->mask_words == 2;
    switch (p->block) {
    case 256:
        return wide ? launch_tiled<OP, PREC, 256, 2>(p, coord, coef, values, prec, singular, s)
                    : launch_tiled<OP, PREC, 256, 1>(p, coord, coef, values, prec, singular, s);
    case 512:
        return wide ? launch_tiled<OP, PREC, 512, 2>(p, coord, coef, values, prec, singular, s)
                    : launch_tiled<OP, PREC, 512, 1>(p, coord, coef, values, prec, singular, s);
    default: return MF_ERR_ARG;
    }
}

struct Resolved {
    int budget, block, max_rows, slot64, slot32;
};

int resolve_params(int op, const mf_plan_params *params, Resolved *r)
{
    const int S = op == MF_OP_ELA ? mf::kStageEla : mf::kStageLap;
    r->budget = 64 * 1024;
    r->block = 256;
    r->max_rows = op == MF_OP_ELA ? 64 : 128;
    if (params) {
        if (params->tile_lds_bytes > 0) r->budget = params->tile_lds_bytes;
        if (params->block_threads > 0) r->block = params->block_threads;
        if (params->max_tile_rows > 0) r->max_rows = params->max_tile_rows;
    }
    r->slot64 = S * 8;
    r->slot32 = S * 4;
    if (r->block != 256 && r->block != 512) return MF_ERR_ARG;
    if (r->budget < 4096 || r->budget > mf::kMaxLdsBytes || r->max_rows > mf::kMaxTileRows)
        return MF_ERR_ARG;
    return MF_OK;
}

int build_checked(mf::PlanHost &h, int op, int nE, int N, int nnz, const int *e2n,
                  const int *row, const int *col, const int *cb, const double *coord,
                  const Resolved &r, int *lds64, int *lds32)
{
    int st = mf::build_plan(h, op, nE, N, nnz, e2n, row, col, cb, coord, r.budget, r.max_rows,
                            r.slot64, r.block);
    if (st != MF_OK) return st;
    st = mf::validate_plan(h, row, col, e2n, r.slot64);
    if (st != MF_OK) return st;
    int m64 = 0, m32 = 0;
    for (const auto &t : h.tiles) {
        m64 = std::max(m64, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, t.n_inc, r.slot64).total);
        m32 = std::max(m32, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, t.n_inc, r.slot32).total);
    }
    if (m64 + 16 > mf::kMaxLdsBytes) return MF_ERR_ARG;  // one row needs more than a CU has
    *lds64 = m64 + 16;
    *lds32 = m32 + 16;
    return MF_OK;
}

// stats: tiles, staged tile elements, CSR entries, incidence entries, plan bytes, LDS
// bytes per workgroup (FP64), block threads, largest tile (rows)
void fill_stats(const mf::PlanHost &h, int lds64, int block, int64_t out[8])
{
    out[0] = (int64_t)h.tiles.size();
    out[1] = (int64_t)h.tile_elems.size();
    out[2] = (int64_t)h.ent_info.size();
    out[3] = (int64_t)h.inc.size();
    out[4] = (int64_t)h.tiles.size() * sizeof(mf::TileHdr) + (int64_t)h.tile_elems.size() * 4 +
             (int64_t)h.tile_verts.size() + (int64_t)h.tile_nodes.size() * 4 +
             (int64_t)h.tile_rowinfo.size() * sizeof(mf::RowInfo) + (int64_t)h.inc.size() * 2 +
             (int64_t)h.ent_mask.size() * 4 + (int64_t)h.ent_info.size() * 2;
    out[5] = lds64;
    out[6] = block;
    out[7] = h.max_rows;
}

bool valid_args(int op, int nE, int N, int nnz, const int *e2n, const int *row, const int *col)
{
    return (op == MF_OP_LAP || op == MF_OP_ELA) && nE >= 0 && N >= 0 && nnz >= 0 &&
           (nE == 0 || e2n) && row && (nnz == 0 || col);
}

}  // namespace

extern "C" int mf_plan_create(mf_plan **plan, int operatorID, int nbElem, int nbNodes, int nnz,
                              const int *elemToNode, const int *row, const int *col,
                              const int *checkBounds, const double *coord,
                              const mf_plan_params *params)
{
    if (!plan) return MF_ERR_ARG;
    *plan = nullptr;
    if (!valid_args(operatorID, nbElem, nbNodes, nnz, elemToNode, row, col)) return MF_ERR_ARG;
    Resolved r;
    int st = resolve_params(operatorID, params, &r);
    if (st != MF_OK) return st;
    mf::PlanHost h;
    int lds64 = 0, lds32 = 0;
    st = build_checked(h, operatorID, nbElem, nbNodes, nnz, elemToNode, row, col, checkBounds,
                       coord, r, &lds64, &lds32);
    if (st != MF_OK) return st;

    mf_plan *p = new (std::nothrow) mf_plan;
    if (!p) return MF_ERR_NOMEM;
    p->op = operatorID;
    p->nE = nbElem;
    p->N = nbNodes;
    p->nnz = nnz;
    p->dim = operatorID == MF_OP_ELA ? 9 : 1;
    p->ntiles = (int)h.tiles.size();
    p->block = r.block;
    p->mask_words = h.mask_words;
    p->lds_fp64 = lds64;
    p->lds_mixed = lds32;
    p->n_isolated = (int)h.isolated.size();
    int rc = upload(&p->tiles, h.tiles);
    if (rc == MF_OK) rc = upload(&p->tile_elems, h.tile_elems);
    if (rc == MF_OK) rc = upload_as(&p->tile_verts, h.tile_verts);
    if (rc == MF_OK) rc = upload(&p->tile_nodes, h.tile_nodes);
    if (rc == MF_OK) rc = upload_as(&p->tile_rowinfo, h.tile_rowinfo);
    if (rc == MF_OK) rc = upload(&p->inc, h.inc);
    if (rc == MF_OK) rc = upload(&p->ent_mask, h.ent_mask);
    if (rc == MF_OK) rc = upload(&p->ent_info, h.ent_info);
    if (rc == MF_OK) rc = upload(&p->isolated, h.isolated);
    if (rc == MF_OK && checkBounds && operatorID == MF_OP_ELA)
        rc = upload(&p->cb, std::vector<int32_t>(checkBounds, checkBounds + (size_t)nbNodes * 3));
    if (rc != MF_OK) {
        free_plan(p);
        return rc;
    }
    fill_stats(h, lds64, r.block, p->stats);
    *plan = p;
    return MF_OK;
}

extern "C" int mf_plan_host_stats(int operatorID, int nbElem, int nbNodes, int nnz,
                                  const int *elemToNode, const int *row, const int *col,
                                  const double *coord, const mf_plan_params *params,
                                  int64_t out[8])
{
    if (!out || !valid_args(operatorID, nbElem, nbNodes, nnz, elemToNode, row, col))
        return MF_ERR_ARG;
    Resolved r;
    int st = resolve_params(operatorID, params, &r);
    if (st != MF_OK) return st;
    mf::PlanHost h;
    int lds64 = 0, lds32 = 0;
    st = build_checked(h, operatorID, nbElem, nbNodes, nnz, elemToNode, row, col, nullptr,
                       coord, r, &lds64, &lds32);
    if (st != MF_OK) return st;
    fill_stats(h, lds64, r.block, out);
    return MF_OK;
}

extern "C" int mf_plan_destroy(mf_plan *plan)
{
    free_plan(plan);
    return MF_OK;
}

extern "C" int mf_plan_stats(const mf_plan *plan, int64_t out[8])
{
    if (!plan || !out) return MF_ERR_ARG;
    std::memcpy(out, plan->stats, sizeof(plan->stats));
    return MF_OK;
}

extern "C" int mf_plan_assemble(mf_plan *plan, const double *coord, const double *coef,
                                double *values, double *prec, int precision, int *singular,
                                void *stream)
{
    if (!plan || !values || !prec || (plan->N > 0 && !coord)) return MF_ERR_ARG;
    if (precision != MF_PREC_FP64 && precision != MF_PREC_MIXED) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (plan->op == MF_OP_LAP)
        return precision == MF_PREC_FP64
                   ? launch_variant<MF_OP_LAP, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s)
                   : launch_variant<MF_OP_LAP, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s);
    return precision == MF_PREC_FP64
               ? launch_variant<MF_OP_ELA, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s)
               : launch_variant<MF_OP_ELA, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s);
}
