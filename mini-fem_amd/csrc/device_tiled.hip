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
    int ntiles = 0, block = 256;
    int lds_fp64 = 0, lds_mixed = 0;
    int n_isolated = 0;
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mf::TileHdr *tiles = nullptr;
    int32_t *tile_elems = nullptr;
    uchar4 *tile_verts = nullptr;
    int32_t *tile_nodes = nullptr;
    mf::RowInfo *tile_rowinfo = nullptr;
    mf::ChunkHdr *chunks = nullptr;
    uint32_t *tasks = nullptr;
    uint16_t *contribs = nullptr;
    int32_t *cb = nullptr;
    int32_t *isolated = nullptr;
};

namespace {

constexpr int kXcd = 8;      // MI355X: workgroups are dealt to the 8 XCDs round-robin
constexpr int kDiagPre = 8;  // diagonal-chunk groups (32 contributions) loaded at once

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

// Sum of one lane's contributions for the G entries packed in `pk` (G = 2 or 4).
template <int OP, typename T, typename T2, int S, int G>
__device__ __forceinline__ void accumulate(const T *__restrict__ stage, const uint32_t (&pk)[G / 2],
                                           double (&acc)[OP == MF_OP_ELA ? 9 : 1])
{
#pragma unroll
    for (int j = 0; j < G; j++) {
        const uint32_t en = (pk[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
        const T *sl = stage + __umul24(en >> 4, S);
        const int a = (en >> 2) & 3, k = en & 3;
        const T2 axy = *reinterpret_cast<const T2 *>(sl + 2 * a);
        const T2 kxy = *reinterpret_cast<const T2 *>(sl + 2 * k);
        const T ga[3] = {axy.x, axy.y, sl[8 + a]};
        const T gk[3] = {kxy.x, kxy.y, sl[8 + k]};
        if constexpr (OP == MF_OP_ELA) {
            const T2 c01 = *reinterpret_cast<const T2 *>(sl + 12);
            T b[9];
            mf::ela_block(ga, gk, c01.x, c01.y, sl[14], b);
#pragma unroll
            for (int q = 0; q < 9; q++) acc[q] += (double)b[q];
        } else {
            const T x = sl[12] * mf::lap_entry(ga, gk);
            acc[0] += (double)x;
        }
    }
}

// One workgroup = one tile. blockIdx -> tile is XCD-aware: XCD x (= blockIdx % 8) walks
// the x-th contiguous eighth of the tiles in partition order, so tiles that share
// elements and nodes run on the same XCD close in time and share its L2.
//
// Global loads are issued in as few dependent rounds as possible: (1) the tile's ids,
// vertex lists, row info and the first diagonal chunk; (2) the coordinate and
// coefficient gathers. gfx950's vmcnt retires loads and stores in order, so a load
// issued after a store waits for that store: every load of a row chunk is issued before
// the previous row chunk's stores, and no other store precedes a load.
template <int OP, int PREC, int BT>
__global__ void __launch_bounds__(BT)
    k_tiled(const mf::TileHdr *__restrict__ tiles, int ntiles,
            const int32_t *__restrict__ tile_elems, const uchar4 *__restrict__ tile_verts,
            const int32_t *__restrict__ tile_nodes, const mf::RowInfo *__restrict__ rowinfo,
            const mf::ChunkHdr *__restrict__ chunks, const uint32_t *__restrict__ tasks,
            const uint16_t *__restrict__ contribs, const double *__restrict__ coord,
            const double *__restrict__ coef, double *__restrict__ values,
            double *__restrict__ prec, int *__restrict__ singular)
{
    using T = typename std::conditional<PREC == MF_PREC_FP64, double, float>::type;
    using T2 = typename Vec2<T>::type;
    constexpr int S = OP == MF_OP_ELA ? mf::kStageEla : mf::kStageLap;
    constexpr int NA = OP == MF_OP_ELA ? 9 : 1;
    constexpr int GR = mf::kRowGroup, GD = mf::kDiagGroup, PF = mf::kRowPrefetch;
    constexpr int NW = BT / 64;
    constexpr int EPT = 8;  // tile elements per thread handled in registers (TE <= EPT*BT)
    extern __shared__ __align__(16) unsigned char lds_raw[];

    const int per = (ntiles + kXcd - 1) / kXcd;
    const int xcd = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    const int tile = xcd * per + slot;
    if (slot >= per || tile >= ntiles) return;
    const mf::TileHdr h = tiles[tile];
    const int NN = h.n_node, TE = h.n_elem, R = h.n_rows;
    const mf::TileLds L = mf::tile_lds_layout(NN, TE, R, S * (int)sizeof(T), NA * 8);

    T *stage = reinterpret_cast<T *>(lds_raw + L.stage);
    double *ldiag = reinterpret_cast<double *>(lds_raw + L.diag);
    double *lprec = reinterpret_cast<double *>(lds_raw + L.prec);
    double *lc = reinterpret_cast<double *>(lds_raw + L.coords);
    int *lrp = reinterpret_cast<int *>(lds_raw + L.rowptr);
    int *lrl = reinterpret_cast<int *>(lds_raw + L.rowlen);
    int *lmask = reinterpret_cast<int *>(lds_raw + L.mask);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t *nodes = tile_nodes + h.node_start;

    // A, round 1: ids, vertex lists, row info (all independent of each other)
    int nodeId[(mf::kMaxTileNodes + BT - 1) / BT];
#pragma unroll
    for (int r = 0; r < (mf::kMaxTileNodes + BT - 1) / BT; r++) {
        const int i = threadIdx.x + r * BT;
        nodeId[r] = i < NN ? nodes[i] : 0;
    }
    int elemId[EPT];
    uchar4 vert[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = threadIdx.x + r * BT;
        elemId[r] = t < TE ? tile_elems[h.elem_start + t] : 0;
        vert[r] = t < TE ? tile_verts[h.elem_start + t] : make_uchar4(0, 0, 0, 0);
    }
    for (int i = threadIdx.x; i < R; i += BT) {
        const mf::RowInfo ri = rowinfo[h.rinfo_start + i];
        lrp[i] = ri.rowptr;
        lrl[i] = ri.len_mask & 0xFFFF;
        lmask[i] = ri.len_mask >> 16;
    }
    // the first diagonal chunk of this wave, all its groups
    uint2 dpre[kDiagPre];
    uint32_t dtask = mf::kNoTask;
    int dgroups = 0;
    const uint2 *dcg = nullptr;
    auto dprefetch = [&](int c) {
        const mf::ChunkHdr chh = chunks[h.dchunk_start + c];
        dtask = tasks[(int64_t)(h.dchunk_start + c) * 64 + lane];
        dcg = reinterpret_cast<const uint2 *>(contribs + chh.contrib_start) + lane;
        dgroups = (chh.steps + GD - 1) / GD;
#pragma unroll
        for (int g = 0; g < kDiagPre; g++) dpre[g] = g < dgroups ? dcg[g * 64] : make_uint2(0u, 0u);
    };
    if (wave < h.n_dchunks) dprefetch(wave);

    // A, round 2: coordinate and coefficient gathers
#pragma unroll
    for (int r = 0; r < (mf::kMaxTileNodes + BT - 1) / BT; r++) {
        const int i = threadIdx.x + r * BT;
        if (i < NN) {
            const double *p = coord + (int64_t)nodeId[r] * 3;
            const double x = p[0], y = p[1], z = p[2];
            lc[i * 3 + 0] = x;
            lc[i * 3 + 1] = y;
            lc[i * 3 + 2] = z;
        }
    }
    double kc[EPT][3];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = threadIdx.x + r * BT;
        kc[r][0] = 1.0;
        kc[r][1] = kc[r][2] = 0.0;
        if (t < TE) {
            if (OP == MF_OP_ELA) {
                const mf::ElaCoef k = mf::ela_coef(coef, elemId[r]);
                kc[r][0] = k.cd;
                kc[r][1] = k.co;
                kc[r][2] = k.cm;
            } else if (coef) {
                kc[r][0] = coef[elemId[r]];
            }
        }
    }
    for (int q = threadIdx.x; q < S; q += BT) stage[TE * S + q] = (T)0;
    __syncthreads();

    // A, compute: scaled gradients + coefficients of every tile element, one slot each
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int t = threadIdx.x + r * BT;
        if (t >= TE) continue;
        const int vi[4] = {vert[r].x, vert[r].y, vert[r].z, vert[r].w};
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
        if (OP == MF_OP_ELA) {
            *reinterpret_cast<T2 *>(sl + 12) = T2{(T)kc[r][0], (T)kc[r][1]};
            *reinterpret_cast<T2 *>(sl + 14) = T2{(T)kc[r][2], (T)0};
            *reinterpret_cast<T2 *>(sl + 16) = T2{(T)0, (T)0};
        } else {
            *reinterpret_cast<T2 *>(sl + 12) = T2{(T)kc[r][0], (T)0};
        }
    }
    __syncthreads();

    // B1: diagonal blocks (balanced chunks), kept in LDS, and their inverses. Padding
    // entries point at the zero slot: they add +0.0, which leaves every sum unchanged.
    for (int ch = wave; ch < h.n_dchunks; ch += NW) {
        double acc[NA];
#pragma unroll
        for (int q = 0; q < NA; q++) acc[q] = 0.0;
#pragma unroll
        for (int g = 0; g < kDiagPre; g++) {
            if (g < dgroups) {
                const uint32_t pk[2] = {dpre[g].x, dpre[g].y};
                accumulate<OP, T, T2, S, GD>(stage, pk, acc);
            }
        }
        for (int g = kDiagPre; g < dgroups; g++) {
            const uint2 v = dcg[g * 64];
            const uint32_t pk[2] = {v.x, v.y};
            accumulate<OP, T, T2, S, GD>(stage, pk, acc);
        }
        const uint32_t w = dtask;
        if (ch + NW < h.n_dchunks) dprefetch(ch + NW);
        if (w == mf::kNoTask) continue;
        const int ra = (int)w;
#pragma unroll
        for (int q = 0; q < NA; q++) ldiag[ra * NA + q] = acc[q];
        if constexpr (OP == MF_OP_ELA) {
            double blk[9];
#pragma unroll
            for (int q = 0; q < 9; q++) blk[q] = acc[q];
            const int info = mf::ela_invert_block(blk, lmask[ra], true);
#pragma unroll
            for (int q = 0; q < 9; q++) lprec[ra * 9 + q] = blk[q];
            if (info != 0 && singular) atomicAdd(singular, 1);
        } else {
            lprec[ra] = 1.0 / acc[0];
        }
    }
    __syncthreads();

    // B2: row chunks. The task word and the first PF contribution groups of the next
    // chunk are loaded before this chunk's stores are issued.
    int ch = wave;
    uint32_t w = 0;
    int ngroups = 0;
    const uint32_t *cg = nullptr;
    uint32_t pre[PF];
    auto prefetch = [&](int c) {
        const mf::ChunkHdr chh = chunks[h.rchunk_start + c];
        w = tasks[(int64_t)(h.rchunk_start + c) * 64 + lane];
        cg = reinterpret_cast<const uint32_t *>(contribs + chh.contrib_start) + lane;
        ngroups = (chh.steps + GR - 1) / GR;
#pragma unroll
        for (int g = 0; g < PF; g++) pre[g] = g < ngroups ? cg[g * 64] : 0u;
    };
    if (ch < h.n_rchunks) prefetch(ch);
    while (ch < h.n_rchunks) {
        double acc[NA];
#pragma unroll
        for (int q = 0; q < NA; q++) acc[q] = 0.0;
#pragma unroll
        for (int g = 0; g < PF; g++) {
            if (g < ngroups) {
                const uint32_t pk[1] = {pre[g]};
                accumulate<OP, T, T2, S, GR>(stage, pk, acc);
            }
        }
        for (int g = PF; g < ngroups; g++) {
            const uint32_t pk[1] = {cg[g * 64]};
            accumulate<OP, T, T2, S, GR>(stage, pk, acc);
        }
        const uint32_t wc = w;
        const int nch = ch + NW;
        if (nch < h.n_rchunks) prefetch(nch);
        if (wc != mf::kNoTask) {
            const int ra = wc & 0xFF, oa = (wc >> 8) & 0xFF;
            if (wc & mf::kDiagLane) {
#pragma unroll
                for (int q = 0; q < NA; q++) acc[q] = ldiag[ra * NA + q];
            }
            double *dst = values + ((int64_t)lrp[ra] + oa) * NA;
#pragma unroll
            for (int q = 0; q < NA; q++) dst[q] = acc[q];
        }
        ch = nch;
    }

    // C: the prec blocks of the tile's rows
    for (int i = wave; i < R; i += NW)
        if (lane < NA && lrl[i] > 0) prec[(int64_t)nodes[i] * NA + lane] = lprec[i * NA + lane];
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
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    if (hipMalloc(reinterpret_cast<void **>(dst), bytes) != hipSuccess) return MF_ERR_NOMEM;
    if (!src.empty() &&
        hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return MF_ERR_HIP;
    return MF_OK;
}

template <typename T>
int upload_raw(T **dst, const T *src, size_t n)
{
    return upload(dst, std::vector<T>(src, src + n));
}

void free_plan(mf_plan *p)
{
    if (!p) return;
    void *ptrs[] = {p->tiles, p->tile_elems, p->tile_verts, p->tile_nodes, p->tile_rowinfo,
                    p->chunks, p->tasks, p->contribs, p->cb, p->isolated};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

template <int OP, int PREC, int BT>
int launch_tiled(mf_plan *p, const double *coord, const double *coef, double *values,
                 double *prec, int *singular, hipStream_t s)
{
    const int lds = PREC == MF_PREC_FP64 ? p->lds_fp64 : p->lds_mixed;
    auto kern = k_tiled<OP, PREC, BT>;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess)
        return MF_ERR_HIP;
    const int per = (p->ntiles + kXcd - 1) / kXcd;
    if (p->ntiles > 0) {
        hipLaunchKernelGGL(kern, dim3(per * kXcd), dim3(BT), lds, s, p->tiles, p->ntiles,
                           p->tile_elems, p->tile_verts, p->tile_nodes, p->tile_rowinfo,
                           p->chunks, p->tasks, p->contribs, coord, coef, values, prec, singular);
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
int launch_bt(mf_plan *p, const double *coord, const double *coef, double *values,
              double *prec, int *singular, hipStream_t s)
{
    switch (p->block) {
    case 128: return launch_tiled<OP, PREC, 128>(p, coord, coef, values, prec, singular, s);
    case 256: return launch_tiled<OP, PREC, 256>(p, coord, coef, values, prec, singular, s);
    case 512: return launch_tiled<OP, PREC, 512>(p, coord, coef, values, prec, singular, s);
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
    if (r->block != 128 && r->block != 256 && r->block != 512) return MF_ERR_ARG;
    if (r->budget < 4096 || r->budget > mf::kMaxLdsBytes || r->max_rows > mf::kMaxTileRows)
        return MF_ERR_ARG;
    return MF_OK;
}

int build_checked(mf::PlanHost &h, int op, int nE, int N, int nnz, const int *e2n,
                  const int *row, const int *col, const int *cb, const double *coord,
                  const Resolved &r, int *lds64, int *lds32)
{
    int st = mf::build_plan(h, op, nE, N, nnz, e2n, row, col, cb, coord, r.budget, r.max_rows,
                            r.slot64);
    if (st != MF_OK) return st;
    st = mf::validate_plan(h, row, col, r.slot64);
    if (st != MF_OK) return st;
    const int bb = op == MF_OP_ELA ? 72 : 8;
    int m64 = 0, m32 = 0;
    for (const auto &t : h.tiles) {
        if (t.n_elem > 8 * r.block) return MF_ERR_ARG;  // the kernel keeps <= 8 per thread
        m64 = std::max(m64, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, r.slot64, bb).total);
        m32 = std::max(m32, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, r.slot32, bb).total);
    }
    if (m64 + 16 > mf::kMaxLdsBytes) return MF_ERR_ARG;  // one row needs more than a CU has
    *lds64 = m64 + 16;
    *lds32 = m32 + 16;
    return MF_OK;
}

void fill_stats(const mf::PlanHost &h, int lds64, int block, int64_t out[8])
{
    out[0] = (int64_t)h.tiles.size();
    out[1] = (int64_t)h.tile_elems.size();
    out[2] = h.total_tasks;
    out[3] = (int64_t)h.contribs.size();
    out[4] = (int64_t)h.tiles.size() * sizeof(mf::TileHdr) + (int64_t)h.tile_elems.size() * 4 +
             (int64_t)h.tile_verts.size() + (int64_t)h.tile_nodes.size() * 4 +
             (int64_t)h.tile_rowinfo.size() * sizeof(mf::RowInfo) +
             (int64_t)h.chunks.size() * sizeof(mf::ChunkHdr) + (int64_t)h.tasks.size() * 4 +
             (int64_t)h.contribs.size() * 2;
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
    p->lds_fp64 = lds64;
    p->lds_mixed = lds32;
    p->n_isolated = (int)h.isolated.size();
    std::vector<uchar4> verts(h.tile_verts.size() / 4);
    std::memcpy(verts.data(), h.tile_verts.data(), h.tile_verts.size());
    int rc = upload(&p->tiles, h.tiles);
    if (rc == MF_OK) rc = upload(&p->tile_elems, h.tile_elems);
    if (rc == MF_OK) rc = upload(&p->tile_verts, verts);
    if (rc == MF_OK) rc = upload(&p->tile_nodes, h.tile_nodes);
    if (rc == MF_OK) rc = upload(&p->tile_rowinfo, h.tile_rowinfo);
    if (rc == MF_OK) rc = upload(&p->chunks, h.chunks);
    if (rc == MF_OK) rc = upload(&p->tasks, h.tasks);
    if (rc == MF_OK) rc = upload(&p->contribs, h.contribs);
    if (rc == MF_OK) rc = upload(&p->isolated, h.isolated);
    if (rc == MF_OK && checkBounds && operatorID == MF_OP_ELA)
        rc = upload_raw(&p->cb, checkBounds, (size_t)nbNodes * 3);
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
                   ? launch_bt<MF_OP_LAP, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s)
                   : launch_bt<MF_OP_LAP, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s);
    return precision == MF_PREC_FP64
               ? launch_bt<MF_OP_ELA, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s)
               : launch_bt<MF_OP_ELA, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s);
}
