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
    int ntiles = 0, block = 512;
    int lds_fp64 = 0, lds_mixed = 0;
    int n_isolated = 0;
    int grid[2] = {0, 0};  // persistent grid per precision, set on its first launch
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mf::TileHdr *tiles = nullptr;
    int32_t *tile_elems = nullptr;
    uint32_t *tile_verts = nullptr;
    int32_t *tile_nodes = nullptr;
    int4 *tile_rowinfo = nullptr;
    uint16_t *inc = nullptr;
    uint32_t *task_info = nullptr;
    int32_t *task_xpos = nullptr;
    uint4 *task_rec = nullptr;
    int32_t *cb = nullptr;
    int32_t *isolated = nullptr;
};

namespace {

constexpr int kXcd = mf::kXcds;
constexpr int kDefaultPairWindow = 0;  // cross-tile pairing off unless asked for
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


// One tile element's gradient of vertex v from its stage slot: {x, y, z}.
// One 3x3 block (9 doubles, 8-byte aligned) as four 16-byte stores and one 8-byte
// store: half the store instructions of nine 8-byte ones, twice the bytes per request
// (gfx950 global stores need only dword alignment).
__device__ __forceinline__ void store_block9(double *dst, const double (&b)[9])
{
#pragma unroll
    for (int q = 0; q < 4; q++)
        *reinterpret_cast<double2 *>(dst + 2 * q) = make_double2(b[2 * q], b[2 * q + 1]);
    dst[8] = b[8];
}

template <typename T, typename T2>
__device__ __forceinline__ void stage_grad(const T *sl, int v, T (&g)[3])
{
    const T2 xy = *reinterpret_cast<const T2 *>(sl + 2 * v);
    g[0] = xy.x;
    g[1] = xy.y;
    g[2] = sl[8 + v];
}

// Persistent workgroups, XCD-aware. The tiles are cut into 8 contiguous ranges in
// partition order, range x for XCD x (workgroup b runs on XCD b % 8); the G workgroups of
// an XCD walk its range together, slot s0, s0 + G, ..., so tiles that share elements and
// nodes run on one XCD close in time and share its L2.
//
// Per tile: fill LDS from registers -> A (element stage) -> B (blocks, stores). The next
// tile's global loads overlap the current tile: its static data (ids, vertex lists, row
// info, incidence lists, task words) is issued after A, its coordinate and coefficient
// gathers after the first block of B has been computed and before that block's stores
// (gfx950's vmcnt retires loads and stores in order, so a load issued behind stores waits
// for them).
template <int OP, int PREC, int BT>
__global__ void __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4)))
    k_tiled(const mf::TileHdr *__restrict__ tiles, int ntiles, int G,
            const int32_t *__restrict__ tile_elems, const uint32_t *__restrict__ tile_verts,
            const int32_t *__restrict__ tile_nodes, const int4 *__restrict__ rowinfo,
            const uint4 *__restrict__ inc8, const uint32_t *__restrict__ task_info,
            const int32_t *__restrict__ task_xpos, const uint4 *__restrict__ task_rec,
            const double *__restrict__ coord,
            const double *__restrict__ coef, double *__restrict__ values,
            double *__restrict__ prec, int *__restrict__ singular)
{
    using T = typename std::conditional<PREC == MF_PREC_FP64, double, float>::type;
    using T2 = typename Vec2<T>::type;
    constexpr bool ELA = OP == MF_OP_ELA;
    constexpr int S = ELA ? mf::kStageEla : mf::kStageLap;
    constexpr int NA = ELA ? 9 : 1;
    constexpr int NW = BT / 64;
    constexpr int EPT = 1;                                     // tile elements per thread
    constexpr int NPT = (mf::kMaxTileNodes + BT - 1) / BT;     // tile nodes per thread
    constexpr int IPT = (mf::kMaxTileInc / 8 + BT - 1) / BT;   // 16-byte incidence loads
    constexpr int PB = mf::kMaxTileBlocks / NW;                // task blocks per wave
    constexpr int TPT = mf::kMaxTileBlocks * 64 / BT;          // task records per thread
    extern __shared__ __align__(16) unsigned char lds_raw[];

    const int per = (ntiles + kXcd - 1) / kXcd;
    const int xcd = blockIdx.x % kXcd;
    int slot = blockIdx.x / kXcd;
    if (slot >= per || xcd * per + slot >= ntiles) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // registers holding one tile's global data: static (round 1) and gathered (round 2)
    int nodeId[NPT], elemId[EPT];
    uint32_t vert[EPT];
    int4 rinf[NPT];
    uint4 incv[IPT];
    uint4 recv[TPT];
    uint32_t pinf[PB];
    int32_t pxp[PB];
    double cx[NPT][3], kc0[EPT], kc1[EPT];

    auto load_static = [&](const mf::TileHdr &hh) {
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const int i = tid + r * BT;
            nodeId[r] = i < hh.n_node ? tile_nodes[hh.node_start + i] : 0;
            rinf[r] = i < hh.n_rows ? rowinfo[hh.rinfo_start + i] : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int t = tid + r * BT;
            elemId[r] = t < hh.n_elem ? tile_elems[hh.elem_start + t] : 0;
            vert[r] = t < hh.n_elem ? tile_verts[hh.elem_start + t] : 0u;
        }
#pragma unroll
        for (int r = 0; r < IPT; r++) {
            const int i = tid + r * BT;
            incv[r] = i < hh.n_inc / 8 ? inc8[hh.inc_start / 8 + i] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < TPT; r++) {
            const int i = tid + r * BT;
            recv[r] = i < hh.n_task ? task_rec[hh.task_start + i] : make_uint4(0, 0, 0, 0);
        }
        int dg, m0, mstep;
        mf::task_blocks(hh.n_rows, hh.n_task, NW, wave, dg, m0, mstep);
#pragma unroll
        for (int k = 0; k < PB; k++) {
            const int idx = (m0 + k * mstep) * 64 + lane;
            pinf[k] = kNoEntry;
            pxp[k] = -1;
            if (m0 >= 0 && idx < hh.n_task) {
                const int64_t e = (int64_t)hh.task_start + idx;
                pinf[k] = task_info[e];
                pxp[k] = task_xpos[e];
            }
        }
    };
    auto load_gathers = [&](const mf::TileHdr &hh) {
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const int i = tid + r * BT;
            cx[r][0] = cx[r][1] = cx[r][2] = 0.0;
            if (i < hh.n_node) {
                const double *p = coord + (int64_t)nodeId[r] * 3;
                cx[r][0] = p[0];
                cx[r][1] = p[1];
                cx[r][2] = p[2];
            }
        }
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int t = tid + r * BT;
            kc0[r] = ELA ? 0.25 : 1.0;  // reference constants: lambda 0.25, mu 1 | k 1
            kc1[r] = 1.0;
            if (t < hh.n_elem && coef) {
                if constexpr (ELA) {
                    const double2 lm = reinterpret_cast<const double2 *>(coef)[elemId[r]];
                    kc0[r] = lm.x;
                    kc1[r] = lm.y;
                } else {
                    kc0[r] = coef[elemId[r]];
                }
            }
        }
    };

    mf::TileHdr h = tiles[xcd * per + slot];
    load_static(h);
    load_gathers(h);
    for (;;) {
        const int NN = h.n_node, TE = h.n_elem, R = h.n_rows;
        const mf::TileLds L =
            mf::tile_lds_layout(NN, TE, R, h.n_inc, h.n_task, S * (int)sizeof(T));
        T *stage = reinterpret_cast<T *>(lds_raw + L.stage);
        int4 *lri = reinterpret_cast<int4 *>(lds_raw + L.rinfo);
        uint4 *ltask = reinterpret_cast<uint4 *>(lds_raw + L.tasks);
        double *lc = reinterpret_cast<double *>(lds_raw + L.coords);
        int *lrnode = reinterpret_cast<int *>(lds_raw + L.rnode);
        uint32_t *lverts = reinterpret_cast<uint32_t *>(lds_raw + L.verts);
        uint16_t *linc = reinterpret_cast<uint16_t *>(lds_raw + L.inc);

        // ---- LDS fill
#pragma unroll
        for (int r = 0; r < NPT; r++) {
            const int i = tid + r * BT;
            if (i < R) {
                lri[i] = rinf[r];
                lrnode[i] = nodeId[r];
            }
            if (i < NN) {
                lc[i * 3 + 0] = cx[r][0];
                lc[i * 3 + 1] = cx[r][1];
                lc[i * 3 + 2] = cx[r][2];
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
        for (int r = 0; r < TPT; r++) {
            const int i = tid + r * BT;
            if (i < h.n_task) ltask[i] = recv[r];
        }
        // the zero slot (index TE): the target of an unpaired second contribution
        if (tid < S) stage[TE * S + tid] = (T)0;
        if (tid == 0) lverts[TE] = 0u;
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
        // this tile's task words, before the registers take the next tile's
        uint32_t tinf[PB];
        int32_t txp[PB];
#pragma unroll
        for (int k = 0; k < PB; k++) {
            tinf[k] = pinf[k];
            txp[k] = pxp[k];
        }
        int dg, m0, mstep;
        mf::task_blocks(R, h.n_task, NW, wave, dg, m0, mstep);
        const int nEB = (h.n_task + 63) / 64;
        __syncthreads();

        // ---- the next tile's static data, in flight during B
        const int nslot = slot + G;
        const bool has_next = nslot < per && xcd * per + nslot < ntiles;
        mf::TileHdr hn = h;
        if (has_next) {
            hn = tiles[xcd * per + nslot];
            load_static(hn);
        }
        bool pending = has_next;  // the next tile's gathers are still to be issued

        // ---- B1: diagonal blocks, one lane per row: the whole incidence list, ascending,
        // four entries per LDS read with their stage reads issued together
        if (dg >= 0) {
            // the diagonal chains are the tile's critical path: this wave's instructions
            // go first on its SIMD while it sums them
            __builtin_amdgcn_s_setprio(3);
            const int r = dg * 64 + lane;
            const bool on = r < R;
            const int4 ri = on ? lri[r] : make_int4(0, 0, 0, 0);
            const int incOff = ri.z & 0xFFFF, val = on ? (ri.z >> 16) & 0xFFFF : 0;
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
                if (pending) {
                    load_gathers(hn);
                    pending = false;
                }
                double blk[9] = {s0, s1, s2, s1, s4, s5, s2, s5, s8};
                if (on) {
                    store_block9(values + dpos * 9, blk);
                }
                const int info = mf::ela_invert_block(blk, (ri.y >> 24) & 7, true);
                if (on) {
                    store_block9(prec + (int64_t)lrnode[r] * 9, blk);
                    if (info != 0 && singular) atomicAdd(singular, 1);
                }
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
                if (pending) {
                    load_gathers(hn);
                    pending = false;
                }
                if (on) {
                    values[dpos] = acc;
                    prec[lrnode[r]] = 1.0 / acc;
                }
            }
        }

        __builtin_amdgcn_s_setprio(0);

        // ---- B2: off-diagonal tasks, one lane per block pair: the elements holding both a
        // and c, two per iteration (independent LDS chains, summed in order); the block
        // goes to row a, its transpose to row c
#pragma clang loop unroll(disable)
        for (int k = 0; k < PB; k++) {
            const int m = m0 + k * mstep;
            if (m0 < 0 || m >= nEB) break;
            const uint32_t inf = tinf[0];
            const int32_t xp = txp[0];
#pragma unroll
            for (int q = 0; q + 1 < PB; q++) {
                tinf[q] = tinf[q + 1];
                txp[q] = txp[q + 1];
            }
            const bool on = inf != kNoEntry;  // lanes past the tile's last task idle
            const int ra = inf & 0xFF, cnt = on ? (inf >> 8) & 0xFF : 0, oa = inf >> 16;
            const int4 ri = on ? lri[ra] : make_int4(0, 0, 0, 0);
            const uint4 rec = on ? ltask[m * 64 + lane] : make_uint4(0, 0, 0, 0);
            double acc[NA];
#pragma unroll
            for (int q = 0; q < NA; q++) acc[q] = 0.0;
            // one contribution pair (element slot, vertex slots of a and c); a padding
            // entry reads the all-zero slot TE, whose products are +0.0 and leave the
            // sums unchanged
            auto add_pair = [&](const int (&t)[2], const int (&a)[2], const int (&kk)[2]) {
                T ga[2][3], gk[2][3], c0[2], c1[2], c2[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const T *sl = stage + t[u] * S;
                    stage_grad<T, T2>(sl, a[u], ga[u]);
                    stage_grad<T, T2>(sl, kk[u], gk[u]);
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
            };
            if (cnt <= mf::kListLen) {
                const uint32_t w[4] = {rec.x, rec.y, rec.z, rec.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (2 * u >= cnt) break;
                    const int e0 = (int)(w[u] & 0xFFFF), e1 = (int)(w[u] >> 16);
                    const int t[2] = {e0 >> 4, e1 >> 4}, a[2] = {(e0 >> 2) & 3, (e1 >> 2) & 3},
                              kk[2] = {e0 & 3, e1 & 3};
                    add_pair(t, a, kk);
                }
            } else {
                // more than kListLen elements: walk the mask over row a's incidence list
                uint64_t msk = (uint64_t)rec.x | (uint64_t)rec.y << 32;
                const int cl = (int)rec.z, incOff = ri.z & 0xFFFF;
                while (msk) {
                    const int b0 = __builtin_ctzll(msk);
                    msk &= msk - 1;
                    const bool two = msk != 0;
                    const int b1 = two ? __builtin_ctzll(msk) : b0;
                    if (two) msk &= msk - 1;
                    const int e0 = linc[incOff + b0], e1 = linc[incOff + b1];
                    const int t[2] = {e0 >> 2, two ? e1 >> 2 : TE}, a[2] = {e0 & 3, e1 & 3};
                    int kk[2];
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        const uint32_t vv = lverts[t[u]];
                        kk[u] = (int)(vv & 0xFF) == cl ? 0
                                : (int)((vv >> 8) & 0xFF) == cl ? 1
                                : (int)((vv >> 16) & 0xFF) == cl ? 2 : 3;
                    }
                    add_pair(t, a, kk);
                }
            }
            if (pending) {
                load_gathers(hn);
                pending = false;
            }
            if (on) {
                if constexpr (ELA) {
                    store_block9(values + ((int64_t)ri.x + oa) * 9, acc);
                    if (xp >= 0) {
                        const double tr[9] = {acc[0], acc[3], acc[6], acc[1], acc[4],
                                              acc[7], acc[2], acc[5], acc[8]};  // transpose
                        store_block9(values + (int64_t)xp * 9, tr);
                    }
                } else {
                    values[(int64_t)ri.x + oa] = acc[0];
                    if (xp >= 0) values[xp] = acc[0];
                }
            }
        }
        if (pending) load_gathers(hn);
        __syncthreads();  // every wave is done with this tile's LDS
        if (!has_next) break;
        h = hn;
        slot = nslot;
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
                    p->inc,   p->task_info,  p->task_xpos,  p->task_rec,   p->cb,
                    p->isolated};
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
    if (p->ntiles > 0) {
        int &G = p->grid[PREC == MF_PREC_FP64 ? 0 : 1];
        if (G == 0) {
            // first launch: allow the whole LDS for this kernel (the attribute is per
            // kernel, shared by every plan), then the persistent grid = as many workgroups
            // per XCD as fit at once, at most one per slot of its tile range. Later
            // launches (also under graph capture) make no runtime queries.
            int dev = 0, cus = 0, per_cu = 0;
            if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    mf::kMaxLdsBytes) != hipSuccess ||
                hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, reinterpret_cast<const void *>(kern), BT, lds) != hipSuccess)
                return MF_ERR_HIP;
            const int per = (p->ntiles + kXcd - 1) / kXcd;
            G = std::max(1, std::min(per, std::max(1, per_cu) * std::max(1, cus / kXcd)));
        }
        hipLaunchKernelGGL(kern, dim3(G * kXcd), dim3(BT), lds, s, p->tiles, p->ntiles, G,
                           p->tile_elems, p->tile_verts, p->tile_nodes, p->tile_rowinfo,
                           reinterpret_cast<const uint4 *>(p->inc), p->task_info, p->task_xpos,
                           p->task_rec, coord, coef, values, prec, singular);
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
    switch (p->block) {
    case 512: return launch_tiled<OP, PREC, 512>(p, coord, coef, values, prec, singular, s);
    case 1024: return launch_tiled<OP, PREC, 1024>(p, coord, coef, values, prec, singular, s);
    default: return MF_ERR_ARG;
    }
}

struct Resolved {
    int budget, block, max_rows, slot64, slot32, pair_window;
};

int resolve_params(int op, const mf_plan_params *params, Resolved *r)
{
    const int S = op == MF_OP_ELA ? mf::kStageEla : mf::kStageLap;
    r->budget = 64 * 1024;
    r->block = 512;
    r->max_rows = op == MF_OP_ELA ? 64 : 128;
    r->pair_window = kDefaultPairWindow;
    if (params) {
        if (params->pair_window != 0) r->pair_window = std::max(params->pair_window, 0);
        if (params->tile_lds_bytes > 0) r->budget = params->tile_lds_bytes;
        if (params->block_threads > 0) r->block = params->block_threads;
        if (params->max_tile_rows > 0) r->max_rows = params->max_tile_rows;
    }
    r->slot64 = S * 8;
    r->slot32 = S * 4;
    if (r->block != 512 && r->block != 1024) return MF_ERR_ARG;
    if (r->budget < 4096 || r->budget > mf::kMaxLdsBytes || r->max_rows > mf::kMaxTileRows)
        return MF_ERR_ARG;
    return MF_OK;
}

int build_checked(mf::PlanHost &h, int op, int nE, int N, int nnz, const int *e2n,
                  const int *row, const int *col, const int *cb, const double *coord,
                  const Resolved &r, int *lds64, int *lds32)
{
    int st = mf::build_plan(h, op, nE, N, nnz, e2n, row, col, cb, coord, r.budget, r.max_rows,
                            r.slot64, r.block, r.pair_window);
    if (st != MF_OK) return st;
    st = mf::validate_plan(h, row, col, e2n, r.slot64);
    if (st != MF_OK) return st;
    int m64 = 0, m32 = 0;
    for (const auto &t : h.tiles) {
        m64 = std::max(m64, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, t.n_inc, t.n_task,
                                                r.slot64).total);
        m32 = std::max(m32, mf::tile_lds_layout(t.n_node, t.n_elem, t.n_rows, t.n_inc, t.n_task,
                                                r.slot32).total);
    }
    if (m64 + 16 > mf::kMaxLdsBytes) return MF_ERR_ARG;  // one row needs more than a CU has
    *lds64 = m64 + 16;
    *lds32 = m32 + 16;
    return MF_OK;
}

// stats: tiles, staged tile elements, off-diagonal tasks, incidence entries, plan
// bytes, LDS bytes per workgroup (FP64), block threads, largest tile (rows)
void fill_stats(const mf::PlanHost &h, int lds64, int block, int64_t out[8])
{
    out[0] = (int64_t)h.tiles.size();
    out[1] = (int64_t)h.tile_elems.size();
    out[2] = (int64_t)h.task_info.size();
    out[3] = (int64_t)h.inc.size();
    out[4] = (int64_t)h.tiles.size() * sizeof(mf::TileHdr) + (int64_t)h.tile_elems.size() * 4 +
             (int64_t)h.tile_verts.size() + (int64_t)h.tile_nodes.size() * 4 +
             (int64_t)h.tile_rowinfo.size() * sizeof(mf::RowInfo) + (int64_t)h.inc.size() * 2 +
             (int64_t)h.task_rec.size() * 4 + (int64_t)h.task_info.size() * 4 +
             (int64_t)h.task_xpos.size() * 4;
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
    int rc = upload(&p->tiles, h.tiles);
    if (rc == MF_OK) rc = upload(&p->tile_elems, h.tile_elems);
    if (rc == MF_OK) rc = upload_as(&p->tile_verts, h.tile_verts);
    if (rc == MF_OK) rc = upload(&p->tile_nodes, h.tile_nodes);
    if (rc == MF_OK) rc = upload_as(&p->tile_rowinfo, h.tile_rowinfo);
    if (rc == MF_OK) rc = upload(&p->inc, h.inc);
    if (rc == MF_OK) rc = upload_as(&p->task_rec, h.task_rec);
    if (rc == MF_OK) rc = upload(&p->task_info, h.task_info);
    if (rc == MF_OK) rc = upload(&p->task_xpos, h.task_xpos);
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
