// Fused per-iteration step on a tiled gather plan (schedule described in plan.h):
// assembly of lap/ela values in the COLORING build's per-entry summation order, plus
// the block-Jacobi prec (prec_init + prec_inversion) -- one launch per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "minifem.h"
#include "abi_guard.h"
#include "plan.h"

struct mf_plan {
    int op = 0, nE = 0, N = 0, nnz = 0, dim = 1;
    int ntiles = 0, block = 512;
    int max_node = 0, max_elem = 0;  // the largest tile: sizes the fixed front of its LDS
    int lds_fp64 = 0, lds_mixed = 0;
    int back_fp64 = 0, back_mixed = 0;  // the largest tile's LDS past the fixed front part
    int n_isolated = 0;
    bool float_stage = false;  // tiles sized for the float stage: the mixed modes only
    int grid[3] = {0, 0, 0};  // persistent grid per precision mode, set at creation
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    mf::TileHdr *tiles = nullptr;
    int32_t *tile_elems = nullptr;
    uint32_t *tile_verts = nullptr;
    int32_t *tile_nodes = nullptr;
    int4 *tile_rowinfo = nullptr;
    uint16_t *inc = nullptr;
    uint4 *task_rec = nullptr;
    int32_t *cb = nullptr;
    int32_t *isolated = nullptr;
};

namespace {

constexpr int kXcd = mf::kXcds;

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

// lap stage slot: vertex v's gradient {x, y, z}.
template <typename T, typename T2>
__device__ __forceinline__ void stage_grad(const T *sl, int v, T (&g)[3])
{
    const T2 xy = *reinterpret_cast<const T2 *>(sl + 2 * v);
    g[0] = xy.x;
    g[1] = xy.y;
    g[2] = sl[8 + v];
}

// ela stage slot: vertex v's group {gx, gy, gz, lambda+2mu} at sl + 4v (16-byte aligned).
__device__ __forceinline__ void ela_vertex(const double *p, double (&g)[3], double &cd)
{
    const double2 a = *reinterpret_cast<const double2 *>(p), b = *reinterpret_cast<const double2 *>(p + 2);
    g[0] = a.x;
    g[1] = a.y;
    g[2] = b.x;
    cd = b.y;
}
__device__ __forceinline__ void ela_vertex(const float *p, float (&g)[3], float &cd)
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    g[0] = a.x;
    g[1] = a.y;
    g[2] = a.z;
    cd = a.w;
}
__device__ __forceinline__ void ela_vertex_store(double *p, const double (&g)[3], double cd)
{
    *reinterpret_cast<double2 *>(p) = make_double2(g[0], g[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(g[2], cd);
}
__device__ __forceinline__ void ela_vertex_store(float *p, const double (&g)[3], double cd)
{
    *reinterpret_cast<float4 *>(p) = make_float4((float)g[0], (float)g[1], (float)g[2], (float)cd);
}

constexpr int kOpsAny = -1;  // operator per job, read from the job at run time

// One plan's launch: its schedule arrays, the buffers its iteration reads and writes, and
// its tile count.
struct TileJob {
    const mf::TileHdr *tiles;
    const int32_t *tile_elems;
    const uint32_t *tile_verts;
    const int32_t *tile_nodes;
    const int4 *rowinfo;
    const uint4 *inc8;
    const uint4 *task_rec;
    const double *coord;
    const double *coef;
    double *values;
    double *prec;
    int *singular;
    int op, ntiles, pad0, pad1;
};

// Several plans in one launch, passed by value in the kernel arguments (no device table
// to upload; a captured graph keeps its own copy).
constexpr int kMaxJobs = MF_MAX_PLAN_JOBS;
struct JobTable {
    TileJob job[kMaxJobs];
    int njobs;
};

// The jobs of a launch as the kernel body sees them.
struct OneJob {
    const TileJob &j;
    __device__ int count() const { return 1; }
    __device__ const TileJob &at(int) const { return j; }
};
struct ManyJobs {
    const JobTable &t;
    __device__ int count() const { return t.njobs; }
    __device__ const TileJob &at(int i) const { return t.job[i]; }
};

// Tile header i, read through the constant address space: a uniform load there is a
// scalar load, which does not queue behind the wave's own stores (a vector load would: the
// compute waves would wait for a tile's stores to drain before starting the next tile).
// Headers are never written during a launch.
__device__ __forceinline__ mf::TileHdr load_hdr(const mf::TileHdr *tiles, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstHdr = const __attribute__((address_space(4))) mf::TileHdr;
    const ConstHdr *p = (const ConstHdr *)(tiles + i);
    mf::TileHdr h;
    h.elem_start = p->elem_start;
    h.n_elem = p->n_elem;
    h.node_start = p->node_start;
    h.n_node = p->n_node;
    h.n_rows = p->n_rows;
    h.rinfo_start = p->rinfo_start;
    h.inc_start = p->inc_start;
    h.n_inc = p->n_inc;
    h.task_start = p->task_start;
    h.n_task = p->n_task;
    h.n_ent = p->n_ent;
    h.pad = 0;
    return h;
#else
    return tiles[i];  // the host pass only parses device code
#endif
}

// The tiles one XCD walks: each job's tiles cut into 8 contiguous ranges, range x for XCD
// x, the ranges of XCD x one job after another. Every XCD thus gets an eighth of every job
// (a launch that mixes cheap lap tiles with costly ela tiles stays balanced), and tiles that
// share elements stay on one XCD. One job: XCD x walks tiles x*ceil(n/8) ... as it always
// has.
template <class Jobs>
struct XcdWalk {
    const Jobs &J;
    int xcd;
    __device__ int per(int j) const { return (J.at(j).ntiles + kXcd - 1) / kXcd; }
    __device__ int len(int j) const
    {
        const int n = J.at(j).ntiles, p = per(j);
        return n - xcd * p < 0 ? 0 : (n - xcd * p < p ? n - xcd * p : p);
    }
    __device__ int total() const
    {
        int L = 0;
        for (int j = 0; j < J.count(); j++) L += len(j);
        return L;
    }
};

// A position of an XCD's walk: the job it falls in and where that job's range begins.
struct Cursor {
    int j, base;
};

// Moves c to position p (positions never decrease along a walk) and returns the tile's
// index in job c.j.
template <class Jobs>
__device__ __forceinline__ int seek(const XcdWalk<Jobs> &W, Cursor &c, int p)
{
    while (p >= c.base + W.len(c.j)) {
        c.base += W.len(c.j);
        c.j++;
    }
    return W.xcd * W.per(c.j) + p - c.base;
}

// LDS layout of tile hh as operator ELA runs it. FRONT_ELA: the fixed front part holds
// ela coefficients (16 bytes per element); a launch that mixes operators sizes it so for
// every tile, since the loaders fill the next tile's front while the current one runs.
template <bool ELA, int PREC, bool FRONT_ELA>
__device__ __forceinline__ mf::TileLds tile_layout(const mf::TileHdr &hh, int f_node, int f_elem)
{
    using T = typename std::conditional<PREC == MF_PREC_FP64, double, float>::type;
    return mf::tile_lds_layout(FRONT_ELA, ELA, f_node, f_elem, hh.n_elem, hh.n_rows, hh.n_inc,
                               hh.n_task, hh.n_ent,
                               mf::stage_len(ELA, PREC == MF_PREC_FP64) * (int)sizeof(T));
}

// One tile on the compute waves (phases A, B, O, W and barriers 1-4), operator ELA; jb
// holds the tile's output buffers. The loader waves meet the same four barriers.
template <bool ELA, int PREC, int BT, bool FRONT_ELA>
__device__ __forceinline__ void compute_tile(const mf::TileHdr &h, const TileJob &jb, int f_node,
                                             int f_elem, unsigned char *lds_raw)
{
    using T = typename std::conditional<PREC == MF_PREC_FP64, double, float>::type;
    using T2 = typename Vec2<T>::type;
    constexpr int S = mf::stage_len(ELA, PREC == MF_PREC_FP64);
    constexpr int NA = ELA ? 9 : 1;
    constexpr int NC = mf::compute_waves(BT), CT = NC * 64;
    constexpr int MAXE = mf::max_tile_elems(BT);
    constexpr int EA = (MAXE + CT - 1) / CT;  // A: elements per thread
    constexpr int TB = mf::task_blocks_per_wave(ELA);  // task blocks per wave
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int NN = h.n_node, TE = h.n_elem, R = h.n_rows;
    (void)NN;
    const mf::TileLds L = tile_layout<ELA, PREC, FRONT_ELA>(h, f_node, f_elem);
    T *stage = reinterpret_cast<T *>(lds_raw + L.stage);
    double *outv = reinterpret_cast<double *>(lds_raw + L.stage);  // the output image
    const int4 *lri = reinterpret_cast<const int4 *>(lds_raw + L.rinfo);
    const uint4 *ltask = reinterpret_cast<const uint4 *>(lds_raw + L.tasks);
    const double *lc = reinterpret_cast<const double *>(lds_raw + L.coords);
    const int *lrnode = reinterpret_cast<const int *>(lds_raw + L.rnode);
    const uint32_t *lverts = reinterpret_cast<const uint32_t *>(lds_raw + L.verts);
    const uint16_t *linc = reinterpret_cast<const uint16_t *>(lds_raw + L.inc);

    // ---- A: scaled gradients + coefficients of every tile element, one slot each
#pragma unroll
    for (int k = 0; k < EA; k++) {
        const int t = tid + k * CT;
        if (t >= TE) continue;
        const uint32_t vv = lverts[t];
        const int vi[4] = {(int)(vv & 0xFF), (int)((vv >> 8) & 0xFF), (int)((vv >> 16) & 0xFF),
                           (int)(vv >> 24)};
        double c[4][3], g[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int d = 0; d < 3; d++) c[i][d] = lc[vi[i] * 3 + d];
        mf::elem_grad(c, g);
        T *sl = stage + t * S;
        if constexpr (ELA) {
            // src/assembly.cc:386-409 with per-element (lambda, mu)
            const double2 lm = reinterpret_cast<const double2 *>(lds_raw + L.coef)[t];
            const double cd = lm.x + 2.0 * lm.y, co = lm.x + lm.y, cm = lm.y;
#pragma unroll
            for (int i = 0; i < 4; i++) ela_vertex_store(sl + 4 * i, g[i], cd);
            *reinterpret_cast<T2 *>(sl + 16) = T2{(T)co, (T)cm};
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                *reinterpret_cast<T2 *>(sl + 2 * i) = T2{(T)g[i][0], (T)g[i][1]};
                sl[8 + i] = (T)g[i][2];
            }
            const double kk = reinterpret_cast<const double *>(lds_raw + L.coef)[t];
            *reinterpret_cast<T2 *>(sl + 12) = T2{(T)kk, (T)0};
        }
    }
    // the zero slot (index TE): the target of an unpaired second contribution
    if (tid < S) stage[TE * S + tid] = (T)0;
    __syncthreads();  // 1

    // ---- B: one work block per compute wave, results held in registers
    const int nDB = mf::diag_blocks(ELA, R), nEB = (h.n_task + 63) / 64;
    const bool isDiag = wave < nDB;
    const int r = wave * 64 + lane;          // diagonal: the lane's row
    const bool dOn = isDiag && r < R;
    int4 dri = make_int4(0, 0, 0, 0);
    double blk[NA];
#pragma unroll
    for (int q = 0; q < NA; q++) blk[q] = 0.0;
    // task blocks of this wave: (wave - nDB) + k * (NC - nDB), k < TB
    bool tOn[TB];
    int tEnt[TB], xEnt[TB];  // tile entries of the task's block and of its transpose (-1: none)
    double tres[TB][NA];
#pragma unroll
    for (int k = 0; k < TB; k++) {
        const int m = (wave - nDB) + k * (NC - nDB);
        tOn[k] = !isDiag && m < nEB && m * 64 + lane < h.n_task;
        tEnt[k] = 0;
        xEnt[k] = -1;
#pragma unroll
        for (int q = 0; q < NA; q++) tres[k][q] = 0.0;
    }
    if (isDiag) {
        // the diagonal chains are the tile's critical path: this wave's instructions
        // go first on its SIMD while it sums them
        __builtin_amdgcn_s_setprio(3);
        dri = dOn ? lri[r] : make_int4(0, 0, 0, 0);
        const int incOff = dri.z & 0xFFFF, val = dOn ? (dri.z >> 16) & 0xFFFF : 0;
        // The row's incidence list, two entries per 32-bit word (lists are padded to a
        // multiple of 4 entries), summed in order one pair at a time. Software
        // pipelined: the LDS reads of the next pair are in flight while the current
        // pair is summed; reads past the list repeat its last word and are never
        // summed.
        const uint32_t *lw = reinterpret_cast<const uint32_t *>(linc + incOff);
        const int np = (val + 1) >> 1, lastw = val > 0 ? ((val + 3) >> 2) * 2 - 1 : 0;
        auto word = [&](int q) { return lw[q < lastw ? q : lastw]; };
        struct Pair {
            T g[2][3];
            T c[2][3];  // ela: lambda + 2 mu, lambda + mu, mu; lap: k
        };
        auto fetch = [&](uint32_t w, Pair &P) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int e = u ? (int)(w >> 16) : (int)(w & 0xFFFF);
                const T *sl = stage + (e >> 2) * S;
                if constexpr (ELA) {
                    ela_vertex(sl + 4 * (e & 3), P.g[u], P.c[u][0]);
                    const T2 c12 = *reinterpret_cast<const T2 *>(sl + 16);
                    P.c[u][1] = c12.x;
                    P.c[u][2] = c12.y;
                } else {
                    stage_grad<T, T2>(sl, e & 3, P.g[u]);
                    P.c[u][0] = sl[12];
                }
            }
        };
        // ela_block(g, g) is exactly symmetric (g_d*g_e == g_e*g_d in IEEE): six
        // sums hold the block bit for bit
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s4 = 0.0, s5 = 0.0, s8 = 0.0;
        auto sum = [&](const Pair &P, int p) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (2 * p + u >= val) break;
                const T *g = P.g[u];
                if constexpr (ELA) {
                    const T cd = P.c[u][0], co = P.c[u][1], cm = P.c[u][2];
                    const T p00 = g[0] * g[0], p11 = g[1] * g[1], p22 = g[2] * g[2];
                    s0 += (double)(p00 * cd + p11 * cm + p22 * cm);
                    s1 += (double)(g[0] * g[1] * co);
                    s2 += (double)(g[0] * g[2] * co);
                    s4 += (double)(p00 * cm + p11 * cd + p22 * cm);
                    s5 += (double)(g[1] * g[2] * co);
                    s8 += (double)(p00 * cm + p11 * cm + p22 * cd);
                } else {
                    const T x = P.c[u][0] * mf::lap_entry(g, g);
                    s0 += (double)x;
                }
            }
        };
        // LDS reads complete in issue order: each pair's list word is read before the
        // previous pair's data, so waiting for the word leaves that data in flight
        Pair A, B;
        uint32_t wa = word(0), wb = word(1);
        fetch(wa, A);  // pair 0
        for (int p = 0; p < np; p += 2) {
            wa = word(p + 2);
            fetch(wb, B);  // pair p + 1
            sum(A, p);
            if (p + 1 >= np) break;
            wb = word(p + 3);
            fetch(wa, A);  // pair p + 2
            sum(B, p + 1);
        }
        if constexpr (ELA) {
            blk[0] = s0; blk[1] = s1; blk[2] = s2;
            blk[3] = s1; blk[4] = s4; blk[5] = s5;
            blk[6] = s2; blk[7] = s5; blk[8] = s8;
        } else {
            blk[0] = s0;
        }
        __builtin_amdgcn_s_setprio(0);
    } else {
#pragma unroll
    for (int kb = 0; kb < TB; kb++) {
        if (!tOn[kb]) continue;
        // off-diagonal task: the elements holding both a and c, two per iteration
        // (independent LDS chains, summed in order); the block goes to row a, its
        // transpose to row c
        const int ti = ((wave - nDB) + kb * (NC - nDB)) * 64 + lane;
        double (&acc)[NA] = tres[kb];
        const uint4 rc = ltask[ti];
        const int ra = rc.w & 0xFF, oa = (rc.w >> 8) & 0xFF, cl = (rc.w >> 16) & 0xFF;
        const int4 ri = lri[ra];
        tEnt[kb] = ri.w + oa;
        if (cl < R) xEnt[kb] = lri[cl].w + (int)(rc.w >> 24);  // c is a row of this tile
        // one contribution pair (element slot, vertex slots of a and c); a padding
        // entry reads the all-zero slot TE, whose products are +0.0 and leave the
        // sums unchanged
        auto add_pair = [&](const int (&t)[2], const int (&a)[2], const int (&kk)[2]) {
            T ga[2][3], gk[2][3], c0[2], c1[2], c2[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const T *sl = stage + t[u] * S;
                if constexpr (ELA) {
                    T unused;
                    ela_vertex(sl + 4 * a[u], ga[u], c0[u]);
                    ela_vertex(sl + 4 * kk[u], gk[u], unused);
                    const T2 c12 = *reinterpret_cast<const T2 *>(sl + 16);
                    c1[u] = c12.x;
                    c2[u] = c12.y;
                } else {
                    stage_grad<T, T2>(sl, a[u], ga[u]);
                    stage_grad<T, T2>(sl, kk[u], gk[u]);
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
        if (rc.x == mf::kWalkRecord) {
            // more than kListLen elements on a row longer than a mask: walk row a's
            // whole incidence list, taking the elements that hold c, two at a time
            const int incOff = ri.z & 0xFFFF, val = (ri.z >> 16) & 0xFFFF;
            int pt = -1, pa = 0, pk = 0;  // a contribution waiting for its pair
            for (int j = 0; j < val; j++) {
                const int e = linc[incOff + j], te = e >> 2;
                const uint32_t vv = lverts[te];
                const int kc = (int)(vv & 0xFF) == cl ? 0
                               : (int)((vv >> 8) & 0xFF) == cl ? 1
                               : (int)((vv >> 16) & 0xFF) == cl ? 2
                               : (int)(vv >> 24) == cl ? 3 : -1;
                if (kc < 0) continue;
                if (pt < 0) {
                    pt = te;
                    pa = e & 3;
                    pk = kc;
                    continue;
                }
                const int t[2] = {pt, te}, a[2] = {pa, e & 3}, kk[2] = {pk, kc};
                add_pair(t, a, kk);
                pt = -1;
            }
            if (pt >= 0) {  // an odd count: the last one with the zero slot
                const int t[2] = {pt, TE}, a[2] = {pa, 0}, kk[2] = {pk, 0};
                add_pair(t, a, kk);
            }
        } else if (rc.x != mf::kMaskRecord) {
            // up to kListLen contributions, padded with the zero slot TE
            const uint32_t w[3] = {rc.x, rc.y, rc.z};
#pragma unroll
            for (int u = 0; u < 3; u++) {
                const int e0 = (int)(w[u] & 0xFFFF), e1 = (int)(w[u] >> 16);
                if ((e0 >> 4) == TE) break;
                const int t[2] = {e0 >> 4, e1 >> 4}, a[2] = {(e0 >> 2) & 3, (e1 >> 2) & 3},
                          kk[2] = {e0 & 3, e1 & 3};
                add_pair(t, a, kk);
            }
        } else {
            // more than kListLen elements: walk the mask over row a's incidence list
            uint64_t msk = (uint64_t)rc.y | (uint64_t)rc.z << 32;
            const int incOff = ri.z & 0xFFFF;
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
    }
    }
    __syncthreads();  // 2: every stage read is done

    // ---- O: the blocks (and in-tile transposes) into the output image
    if (dOn) {
        double *o = outv + (int64_t)(dri.w + ((dri.y >> 16) & 0xFF)) * NA;
#pragma unroll
        for (int q = 0; q < NA; q++) o[q] = blk[q];
    }
#pragma unroll
    for (int kb = 0; kb < TB; kb++) {
        if (!tOn[kb]) continue;
        const double (&r)[NA] = tres[kb];
        double *o = outv + (int64_t)tEnt[kb] * NA;
#pragma unroll
        for (int q = 0; q < NA; q++) o[q] = r[q];
        if (xEnt[kb] >= 0) {
            if constexpr (ELA) {
                const double tr[9] = {r[0], r[3], r[6], r[1], r[4],
                                      r[7], r[2], r[5], r[8]};  // transpose
                double *ot = outv + (int64_t)xEnt[kb] * 9;
#pragma unroll
                for (int q = 0; q < 9; q++) ot[q] = tr[q];
            } else {
                outv[xEnt[kb]] = r[0];
            }
        }
    }
    __syncthreads();  // 3

    // ---- W: the diagonal waves invert their blocks into prec; the other compute
    // waves copy the image to the rows' CSR segments (each segment contiguous in both)
    if (isDiag) {
        if (dOn) {
            const int node = lrnode[r];
            if constexpr (ELA) {
                const int info = mf::ela_invert_block(blk, (dri.y >> 24) & 7, true);
                store_block9(jb.prec + (int64_t)node * 9, blk);
                if (info != 0 && jb.singular) atomicAdd(jb.singular, 1);
            } else {
                jb.prec[node] = 1.0 / blk[0];
            }
        }
    } else {
        const int nwr = NC - nDB, wr = wave - nDB;
        if constexpr (ELA) {
            // four rows per wave at a time, 16 lanes each. A row is 9 len doubles
            // (~135) moved as 16-byte pieces, one 8-byte piece at the end of an
            // odd-length row (8-byte aligned 16-byte stores are valid on gfx950; the
            // LDS side reads two 8-byte words). A lane reads all its pieces of a pass
            // (CP of them, 32 CP doubles per row and pass) before it stores any, so
            // a row costs one LDS round trip, not one per piece.
            constexpr int CP = 5;
            const int g = lane >> 4, l16 = lane & 15;
            for (int rr = wr * 4 + g; rr < R; rr += nwr * 4) {
                const int4 ri = lri[rr];
                const int n = (ri.y & 0xFFFF) * 9;
                const double *src = outv + (int64_t)ri.w * 9;
                double *dst = jb.values + (int64_t)ri.x * 9;
                for (int j0 = 2 * l16; j0 < n; j0 += 32 * CP) {
                    double2 v[CP];
#pragma unroll
                    for (int k = 0; k < CP; k++) {
                        const int j = j0 + 32 * k;
                        v[k] = make_double2(j < n ? src[j] : 0.0, j + 1 < n ? src[j + 1] : 0.0);
                    }
#pragma unroll
                    for (int k = 0; k < CP; k++) {
                        const int j = j0 + 32 * k;
                        if (j + 1 < n)
                            *reinterpret_cast<double2 *>(dst + j) = v[k];
                        else if (j < n)
                            dst[j] = v[k].x;
                    }
                }
            }
        } else {
            // four rows per wave, 16 lanes each
            const int g = lane >> 4, l16 = lane & 15;
            for (int rr = wr * 4 + g; rr < R; rr += nwr * 4) {
                const int4 ri = lri[rr];
                const int n = ri.y & 0xFFFF;
                for (int j = l16; j < n; j += 16) jb.values[(int64_t)ri.x + j] = outv[ri.w + j];
            }
        }
    }
    // the next tile's vertices, coordinates and coefficients are written meanwhile
    __syncthreads();  // 4
}

// Persistent workgroups, XCD-aware. The tiles of a launch (one plan's, or the plans of a
// job table one after another) are cut into 8 contiguous ranges in that order, range x
// for XCD x (workgroup b runs on XCD b % 8); the G workgroups of an XCD walk its range
// together, slot s0, s0 + G, ..., so tiles that share elements and nodes run on one XCD
// close in time and share its L2.
//
// Waves 0..NC-1 compute, waves NC.. load. Per tile, separated by four barriers:
//   A  compute: element stage          | load: row data, tasks, incidences into LDS
//   B  compute: blocks into registers  | load: next tile's static loads and gathers
//   O  compute: blocks into the output image (laid over the stage)
//   W  compute: image -> values, prec  | load: wait gathers, next vertices/coords/coefs
// Compute waves issue no global loads and loader waves no stores, so no wave ever waits
// for its own stores to drain.
template <int OPS, int PREC, int BT, class Jobs>
__device__ __forceinline__ void tiled_body(const Jobs &J, int G, int f_node, int f_elem)
{
    constexpr bool FRONT_ELA = OPS != MF_OP_LAP;
    constexpr int NC = mf::compute_waves(BT), CT = NC * 64, LT = BT - CT;
    constexpr int MAXE = mf::max_tile_elems(BT);
    constexpr int NPT = (mf::kMaxTileNodes + LT - 1) / LT;               // loader: nodes, rows
    constexpr int EPT = (MAXE + LT - 1) / LT;                            // loader: elements
    constexpr int IPT = (mf::kMaxTileInc / 8 + LT - 1) / LT;             // loader: 16-byte incidences
    // loader: tasks (a launch that mixes operators holds the larger lap count)
    constexpr int TPT = (mf::max_tile_tasks(OPS == MF_OP_ELA, BT) + LT - 1) / LT;
    static_assert(mf::kMaxTileRows <= mf::kMaxTileNodes, "rows are the first tile nodes");
    extern __shared__ __align__(16) unsigned char lds_raw[];

    const int xcd = blockIdx.x % kXcd;
    const XcdWalk<Jobs> W{J, xcd};
    const int nslots = W.total();  // this XCD's tiles
    int slot = blockIdx.x / kXcd;
    if (slot >= nslots) return;
    const int tid = threadIdx.x, wave = tid >> 6;
    auto is_ela = [&](const TileJob &jb) -> bool {
        if constexpr (OPS == kOpsAny) return jb.op == MF_OP_ELA;
        else return OPS == MF_OP_ELA;
    };
    auto layout = [&](const mf::TileHdr &hh, bool ela) {
        return ela ? tile_layout<true, PREC, FRONT_ELA>(hh, f_node, f_elem)
                   : tile_layout<false, PREC, FRONT_ELA>(hh, f_node, f_elem);
    };
    Cursor cur{0, 0};
    mf::TileHdr h = load_hdr(J.at(0).tiles, 0);
    {
        const int t = seek(W, cur, slot);
        h = load_hdr(J.at(cur.j).tiles, t);
    }
    int jcur = cur.j;

    if (wave >= NC) {
        // ------------------------------------------------------------------ loader waves
        const int lt = tid - CT;
        // the node and element ids of a tile are loaded one tile earlier than the rest of
        // its data, so its coordinate and coefficient gathers go out together with its
        // static loads: one memory round trip per tile
        int nodeId[NPT], nodeIdN[NPT];
        int elemId[EPT], elemIdN[EPT];
        double cx[NPT][3];
        // 16-byte records as four scalar arrays each (an array of vectors stays in scratch)
        int rix[NPT], riy[NPT], riz[NPT], riw[NPT];
        uint32_t vert[EPT];
        double k0[EPT], k1[EPT];  // raw coefficients: ela (lambda, mu), lap (k, unused)
        uint32_t inx[IPT], iny[IPT], inz[IPT], inw[IPT];
        uint32_t rcx[TPT], rcy[TPT], rcz[TPT], rcw[TPT];
        // every load is unconditional, from an index clamped into the tile's range (the
        // fill skips the lanes past it): no branch between the loads, so they all stay in
        // flight together
        auto at = [](int start, int n, int i) -> int64_t {
            return n > 0 ? (int64_t)start + (i < n ? i : n - 1) : 0;
        };
        auto load_ids = [&](const TileJob &jb, const mf::TileHdr &hh, int (&nid)[NPT],
                            int (&eid)[EPT]) {
#pragma unroll
            for (int r = 0; r < NPT; r++)
                nid[r] = jb.tile_nodes[at(hh.node_start, hh.n_node, lt + r * LT)];
#pragma unroll
            for (int r = 0; r < EPT; r++)
                eid[r] = jb.tile_elems[at(hh.elem_start, hh.n_elem, lt + r * LT)];
        };
        auto load_static = [&](const TileJob &jb, const mf::TileHdr &hh) {
#pragma unroll
            for (int r = 0; r < NPT; r++) {
                const int i = lt + r * LT;
                const int4 v = jb.rowinfo[at(hh.rinfo_start, hh.n_rows, i)];
                rix[r] = v.x;
                riy[r] = v.y;
                riz[r] = v.z;
                riw[r] = v.w;
            }
#pragma unroll
            for (int r = 0; r < EPT; r++) {
                const int i = lt + r * LT;
                vert[r] = jb.tile_verts[at(hh.elem_start, hh.n_elem, i)];
            }
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const uint4 v = jb.inc8[at(hh.inc_start / 8, hh.n_inc / 8, lt + r * LT)];
                inx[r] = v.x;
                iny[r] = v.y;
                inz[r] = v.z;
                inw[r] = v.w;
            }
#pragma unroll
            for (int r = 0; r < TPT; r++) {
                const int64_t e = at(hh.task_start, hh.n_task, lt + r * LT);
                const uint4 v = jb.task_rec[e];
                rcx[r] = v.x;
                rcy[r] = v.y;
                rcz[r] = v.z;
                rcw[r] = v.w;
            }
        };
        auto load_gathers = [&](const TileJob &jb) {
            const double *coord = jb.coord, *coef = jb.coef;
#pragma unroll
            for (int r = 0; r < NPT; r++) {
                const double *p = coord + (int64_t)nodeId[r] * 3;
                cx[r][0] = p[0];
                cx[r][1] = p[1];
                cx[r][2] = p[2];
            }
            // without a coefficient array, a valid dummy address keeps the loads
            // branch-free (fill1 substitutes the constants). The operator branch stays
            // outside the unrolled loops, so k0/k1 are indexed by constants only.
            const bool hc = coef != nullptr;
            if constexpr (PREC == MF_PREC_MIXED_F32COEF) {  // a float coefficient field
                const float *cf = reinterpret_cast<const float *>(hc ? coef : coord);
                if (is_ela(jb)) {
#pragma unroll
                    for (int r = 0; r < EPT; r++) {
                        const float2 lm = reinterpret_cast<const float2 *>(cf)[hc ? elemId[r] : 0];
                        k0[r] = lm.x;
                        k1[r] = lm.y;
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < EPT; r++) {
                        k0[r] = cf[hc ? elemId[r] : 0];
                        k1[r] = 0.0;  // both branches define every register of k0 and k1
                    }
                }
            } else if (is_ela(jb)) {
                const double2 *cf = reinterpret_cast<const double2 *>(hc ? coef : coord);
#pragma unroll
                for (int r = 0; r < EPT; r++) {
                    const double2 lm = cf[hc ? elemId[r] : 0];
                    k0[r] = lm.x;
                    k1[r] = lm.y;
                }
            } else {
                const double *cf = hc ? coef : coord;
#pragma unroll
                for (int r = 0; r < EPT; r++) {
                    k0[r] = cf[hc ? elemId[r] : 0];
                    k1[r] = 0.0;
                }
            }
        };
        // vertices, coordinates, coefficients: what A reads
        auto fill1 = [&](const TileJob &jb, const mf::TileHdr &hh) {
            const bool ela = is_ela(jb);
            const mf::TileLds L = layout(hh, ela);
            double *lc = reinterpret_cast<double *>(lds_raw + L.coords);
            uint32_t *lverts = reinterpret_cast<uint32_t *>(lds_raw + L.verts);
#pragma unroll
            for (int r = 0; r < NPT; r++) {
                const int i = lt + r * LT;
                if (i < hh.n_node) {
                    lc[i * 3 + 0] = cx[r][0];
                    lc[i * 3 + 1] = cx[r][1];
                    lc[i * 3 + 2] = cx[r][2];
                }
            }
            // without a coefficient array: the reference constants (lambda 0.25, mu 1 | k 1)
            const bool hc = jb.coef != nullptr;
#pragma unroll
            for (int r = 0; r < EPT; r++) {
                const int i = lt + r * LT;
                if (i < hh.n_elem) lverts[i] = vert[r];
            }
            if (ela) {
                double2 *lk = reinterpret_cast<double2 *>(lds_raw + L.coef);
#pragma unroll
                for (int r = 0; r < EPT; r++) {
                    const int i = lt + r * LT;
                    if (i < hh.n_elem) lk[i] = make_double2(hc ? k0[r] : 0.25, hc ? k1[r] : 1.0);
                }
            } else {
                double *lk = reinterpret_cast<double *>(lds_raw + L.coef);
#pragma unroll
                for (int r = 0; r < EPT; r++) {
                    const int i = lt + r * LT;
                    if (i < hh.n_elem) lk[i] = hc ? k0[r] : 1.0;
                }
            }
            if (lt == 0) lverts[hh.n_elem] = 0u;  // the zero slot's vertices
        };
        // row info, row nodes, incidence lists, tasks: what B, O and W read
        auto fill2 = [&](const TileJob &jb, const mf::TileHdr &hh) {
            const mf::TileLds L = layout(hh, is_ela(jb));
            int4 *lri = reinterpret_cast<int4 *>(lds_raw + L.rinfo);
            int *lrnode = reinterpret_cast<int *>(lds_raw + L.rnode);
#pragma unroll
            for (int r = 0; r < NPT; r++) {
                const int i = lt + r * LT;
                if (i < hh.n_rows) {
                    lri[i] = make_int4(rix[r], riy[r], riz[r], riw[r]);
                    lrnode[i] = nodeId[r];
                }
            }
#pragma unroll
            for (int r = 0; r < IPT; r++) {
                const int i = lt + r * LT;
                if (i < hh.n_inc / 8)
                    reinterpret_cast<uint4 *>(lds_raw + L.inc)[i] = make_uint4(inx[r], iny[r], inz[r], inw[r]);
            }
#pragma unroll
            for (int r = 0; r < TPT; r++) {
                const int i = lt + r * LT;
                if (i < hh.n_task) {
                    reinterpret_cast<uint4 *>(lds_raw + L.tasks)[i] = make_uint4(rcx[r], rcy[r], rcz[r], rcw[r]);
                }
            }
        };

        auto valid = [&](int sl) { return sl < nslots; };
        // jcur: the job of tile h; cIds, tIds: the walk position and tile index of the tile
        // whose ids are in nodeIdN
        Cursor cIds = cur;
        int tIds = 0;
        load_ids(J.at(jcur), h, nodeId, elemId);
        load_static(J.at(jcur), h);
        load_gathers(J.at(jcur));
        if (valid(slot + G)) {
            tIds = seek(W, cIds, slot + G);
            load_ids(J.at(cIds.j), load_hdr(J.at(cIds.j).tiles, tIds), nodeIdN, elemIdN);
        }
        fill1(J.at(jcur), h);
        __syncthreads();  // P
        for (;;) {
            // A
            fill2(J.at(jcur), h);
            const int nslot = slot + G;
            const bool has_next = valid(nslot);
            if (has_next) {  // the next tile's ids, loaded a tile ago
#pragma unroll
                for (int r = 0; r < NPT; r++) nodeId[r] = nodeIdN[r];
#pragma unroll
                for (int r = 0; r < EPT; r++) elemId[r] = elemIdN[r];
            }
            __syncthreads();  // 1
            // B: the next tile's static data and gathers, and the ids of the one after
            mf::TileHdr hn = h;
            int jn = jcur;
            if (has_next) {
                jn = cIds.j;
                hn = load_hdr(J.at(jn).tiles, tIds);
                load_static(J.at(jn), hn);
                load_gathers(J.at(jn));
                if (valid(nslot + G)) {
                    tIds = seek(W, cIds, nslot + G);
                    load_ids(J.at(cIds.j), load_hdr(J.at(cIds.j).tiles, tIds), nodeIdN, elemIdN);
                }
            }
            __syncthreads();  // 2
            // O
            __syncthreads();  // 3
            // W
            if (has_next) fill1(J.at(jn), hn);
            __syncthreads();  // 4
            if (!has_next) break;
            h = hn;
            jcur = jn;
            slot = nslot;
        }
        return;
    }

    // ---------------------------------------------------------------------- compute waves
    __syncthreads();  // P
    for (;;) {
        const TileJob &jb = J.at(jcur);
        if (is_ela(jb)) compute_tile<OPS != MF_OP_LAP, PREC, BT, FRONT_ELA>(h, jb, f_node, f_elem, lds_raw);
        else compute_tile<OPS == MF_OP_ELA, PREC, BT, FRONT_ELA>(h, jb, f_node, f_elem, lds_raw);
        const int nslot = slot + G;
        if (nslot >= nslots) break;
        slot = nslot;
        const int t = seek(W, cur, slot);
        jcur = cur.j;
        h = load_hdr(J.at(jcur).tiles, t);
    }
}

template <int OP, int PREC, int BT>
__global__ void __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4)))
    k_tiled(TileJob job, int G, int f_node, int f_elem)
{
    tiled_body<OP, PREC, BT>(OneJob{job}, G, f_node, f_elem);
}

// The tiles of up to kMaxJobs plans, lap and ela mixed, in one persistent launch.
template <int PREC, int BT>
__global__ void __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4)))
    k_tiled_jobs(JobTable table, int G, int f_node, int f_elem)
{
    tiled_body<kOpsAny, PREC, BT>(ManyJobs{table}, G, f_node, f_elem);
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

// The plan's node order: the rows of tile 0, tile 1, ... (each tile's rows are spatially
// close), then the nodes with an empty row. Tile t's rows sit at order[rinfo_start ..]:
// the row-info array holds one entry per tile row, in tile order.
__global__ void k_row_order(const mf::TileHdr *__restrict__ tiles, int ntiles,
                            const int32_t *__restrict__ tile_nodes,
                            const int32_t *__restrict__ isolated, int n_isolated, int n_rows,
                            int *__restrict__ order)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const mf::TileHdr h = tiles[t];
        for (int r = threadIdx.x; r < h.n_rows; r += blockDim.x)
            order[h.rinfo_start + r] = tile_nodes[h.node_start + r];
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_isolated; k += gridDim.x * blockDim.x)
        order[n_rows + k] = isolated[k];
}

// Write-only probe of the assembly's output stream. Mode 0: every tile row's CSR segment,
// tiles dealt to XCDs and walked as k_tiled walks them, each segment written front to
// back in 16-byte pieces (the copy-out of phase W); mode 1: the whole value array front to
// back (the same bytes: the difference is what the row order costs in HBM); modes 2 and 3:
// mode 0 split into the 128-byte lines a segment covers whole (2) and the partial lines
// at its two ends, which it shares with the neighbouring rows (3).
template <int D>
__global__ void __launch_bounds__(1024)
    k_write_probe(const mf::TileHdr *__restrict__ tiles, int ntiles, int G,
                  const int4 *__restrict__ rowinfo, double *__restrict__ values, int mode,
                  int64_t total)
{
    if (mode == 1) {
        const int64_t n2 = total / 2, stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
            reinterpret_cast<double2 *>(values)[i] = make_double2(0.0, 0.0);
        if (blockIdx.x == 0 && threadIdx.x == 0 && (total & 1)) values[total - 1] = 0.0;
        return;
    }
    const int per = (ntiles + kXcd - 1) / kXcd, xcd = blockIdx.x % kXcd;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int slot = blockIdx.x / kXcd; slot < per && xcd * per + slot < ntiles; slot += G) {
        const mf::TileHdr h = tiles[xcd * per + slot];
        for (int r = wave; r < h.n_rows; r += nw) {
            const int4 ri = rowinfo[h.rinfo_start + r];
            const int64_t d0 = (int64_t)ri.x * D, d1 = d0 + (int64_t)(ri.y & 0xFFFF) * D;
            if (mode == 0) {
                for (int64_t j = d0 + 2 * lane; j < d1; j += 128) {
                    if (j + 1 < d1) *reinterpret_cast<double2 *>(values + j) = make_double2(0.0, 0.0);
                    else values[j] = 0.0;
                }
                continue;
            }
            // doubles [l0, l1) are the segment's whole 128-byte lines (16 doubles each)
            const int64_t l0 = (d0 + 15) / 16 * 16, l1 = d1 / 16 * 16;
            if (mode == 2) {
                for (int64_t j = l0 + 2 * lane; j < l1; j += 128)
                    *reinterpret_cast<double2 *>(values + j) = make_double2(0.0, 0.0);
            } else {
                const int64_t h1 = l0 < d1 ? l0 : d1, t0 = l1 > h1 ? l1 : h1;
                for (int64_t j = d0 + lane; j < h1; j += 64) values[j] = 0.0;
                for (int64_t j = t0 + lane; j < d1; j += 64) values[j] = 0.0;
            }
        }
    }
}

TileJob make_job(const mf_plan *p, const double *coord, const double *coef, double *values,
                 double *prec, int *singular)
{
    TileJob j{};
    j.tiles = p->tiles;
    j.tile_elems = p->tile_elems;
    j.tile_verts = p->tile_verts;
    j.tile_nodes = p->tile_nodes;
    j.rowinfo = p->tile_rowinfo;
    j.inc8 = reinterpret_cast<const uint4 *>(p->inc);
    j.task_rec = p->task_rec;
    j.coord = coord;
    j.coef = coef;
    j.values = values;
    j.prec = prec;
    j.singular = singular;
    j.op = p->op;
    j.ntiles = p->ntiles;
    return j;
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
                    p->inc,   p->task_rec,   p->cb,
                    p->isolated};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

// At plan creation, for each precision mode: allow the whole LDS for this kernel (the
// attribute is per kernel, shared by every plan), then the persistent grid = as many
// workgroups per XCD as fit at once, at most one per slot of its tile range. Launches
// then only read the plan: they make no runtime queries (also under graph capture), and
// several host threads may launch one plan at once.
template <int OP, int PREC, int BT>
int prepare_tiled(mf_plan *p)
{
    if (p->ntiles == 0) return MF_OK;
    const int lds = PREC == MF_PREC_FP64 ? p->lds_fp64 : p->lds_mixed;
    auto kern = k_tiled<OP, PREC, BT>;
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
    p->grid[PREC] = std::max(1, std::min(per, std::max(1, per_cu) * std::max(1, cus / kXcd)));
    return MF_OK;
}

template <int OP, int PREC, int BT>
int launch_tiled(mf_plan *p, const double *coord, const double *coef, double *values,
                 double *prec, int *singular, hipStream_t s)
{
    const int lds = PREC == MF_PREC_FP64 ? p->lds_fp64 : p->lds_mixed;
    auto kern = k_tiled<OP, PREC, BT>;
    if (p->ntiles > 0) {
        const int G = p->grid[PREC];
        if (G < 1) return MF_ERR_STATE;  // not prepared at creation
        hipLaunchKernelGGL(kern, dim3(G * kXcd), dim3(BT), lds, s,
                           make_job(p, coord, coef, values, prec, singular), G, p->max_node,
                           p->max_elem);
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
int prepare_variant(mf_plan *p)
{
    switch (p->block) {
    case 512: return prepare_tiled<OP, PREC, 512>(p);
    case 1024: return prepare_tiled<OP, PREC, 1024>(p);
    default: return MF_ERR_ARG;
    }
}

template <int OP>
int prepare_all(mf_plan *p)
{
    int st = p->float_stage ? MF_OK : prepare_variant<OP, MF_PREC_FP64>(p);
    if (st == MF_OK) st = prepare_variant<OP, MF_PREC_MIXED>(p);
    if (st == MF_OK) st = prepare_variant<OP, MF_PREC_MIXED_F32COEF>(p);
    return st;
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
    int budget, block, max_rows, slot64, slot32, id_chunks;
    bool float_stage;  // size the tiles by the float stage slot (mixed modes only)
};

int resolve_params(int op, const mf_plan_params *params, Resolved *r)
{
    const bool ela = op == MF_OP_ELA;
    // the fastest schedule measured on MI355X for the lap and ela cubes of 27K-1M nodes
    r->budget = 155000;
    r->block = 1024;
    r->max_rows = 254;
    r->id_chunks = 1;
    r->float_stage = false;
    if (params) {
        if (params->tile_lds_bytes > 0) r->budget = params->tile_lds_bytes;
        if (params->block_threads > 0) r->block = params->block_threads;
        if (params->max_tile_rows > 0) r->max_rows = params->max_tile_rows;
        if (params->id_chunks > 0) r->id_chunks = params->id_chunks;
        const int sp = params->stage_precision;
        if (sp != MF_PREC_FP64 && sp != MF_PREC_MIXED && sp != MF_PREC_MIXED_F32COEF)
            return MF_ERR_ARG;
        r->float_stage = sp != MF_PREC_FP64;
    }
    r->slot64 = mf::stage_len(ela, true) * 8;
    r->slot32 = mf::stage_len(ela, false) * 4;
    if (r->block != 512 && r->block != 1024) return MF_ERR_ARG;
    if (r->budget < 4096 || r->budget > mf::kMaxLdsBytes || r->max_rows > mf::kMaxTileRows)
        return MF_ERR_ARG;
    return MF_OK;
}

int build_checked(mf::PlanHost &h, int op, int nE, int N, int nnz, const int *e2n,
                  const int *row, const int *col, const int *cb, const double *coord,
                  const Resolved &r, int *lds64, int *lds32, int *back64 = nullptr,
                  int *back32 = nullptr)
{
    // the stage slot the tiles are sized by: a float-staged plan's FP64 layout may exceed
    // one CU's LDS, and that plan never launches in FP64
    const int slot = r.float_stage ? r.slot32 : r.slot64;
    int st = mf::build_plan(h, op, nE, N, nnz, e2n, row, col, cb, coord, r.budget, r.max_rows,
                            slot, r.block, r.id_chunks);
    if (st != MF_OK) return st;
    st = mf::validate_plan(h, row, col, e2n, slot);
    if (st != MF_OK) return st;
    int m64 = 0, m32 = 0, b64 = 0, b32 = 0;
    for (const auto &t : h.tiles) {
        // the front part (fixed offsets) is sized by the largest tile
        const bool ela = op == MF_OP_ELA;
        const mf::TileLds L64 = mf::tile_lds_layout(ela, h.max_node, h.max_elem, t.n_elem,
                                                    t.n_rows, t.n_inc, t.n_task, t.n_ent, r.slot64);
        const mf::TileLds L32 = mf::tile_lds_layout(ela, h.max_node, h.max_elem, t.n_elem,
                                                    t.n_rows, t.n_inc, t.n_task, t.n_ent, r.slot32);
        m64 = std::max(m64, L64.total);
        m32 = std::max(m32, L32.total);
        b64 = std::max(b64, L64.total - L64.stage);
        b32 = std::max(b32, L32.total - L32.stage);
    }
    if ((r.float_stage ? m32 : m64) + 16 > mf::kMaxLdsBytes)
        return MF_ERR_ARG;  // one row needs more than a CU has
    *lds64 = m64 + 16;
    *lds32 = m32 + 16;
    if (back64) *back64 = b64;
    if (back32) *back32 = b32;
    return MF_OK;
}

// stats: tiles, staged tile elements, off-diagonal tasks, incidence entries, plan
// bytes, LDS bytes per workgroup (in the precision the tiles are sized for), block
// threads, largest tile (rows)
void fill_stats(const mf::PlanHost &h, int lds64, int block, int64_t out[8])
{
    out[0] = (int64_t)h.tiles.size();
    out[1] = (int64_t)h.tile_elems.size();
    out[2] = (int64_t)h.task_rec.size() / 4;
    out[3] = (int64_t)h.inc.size();
    out[4] = (int64_t)h.tiles.size() * sizeof(mf::TileHdr) + (int64_t)h.tile_elems.size() * 4 +
             (int64_t)h.tile_verts.size() + (int64_t)h.tile_nodes.size() * 4 +
             (int64_t)h.tile_rowinfo.size() * sizeof(mf::RowInfo) + (int64_t)h.inc.size() * 2 +
             (int64_t)h.task_rec.size() * 4;
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
    return mf_guarded([&]() -> int {
        if (!plan) return MF_ERR_ARG;
        *plan = nullptr;
        if (!valid_args(operatorID, nbElem, nbNodes, nnz, elemToNode, row, col)) return MF_ERR_ARG;
        Resolved r;
        int st = resolve_params(operatorID, params, &r);
        if (st != MF_OK) return st;
        mf::PlanHost h;
        int lds64 = 0, lds32 = 0, back64 = 0, back32 = 0;
        st = build_checked(h, operatorID, nbElem, nbNodes, nnz, elemToNode, row, col, checkBounds,
                           coord, r, &lds64, &lds32, &back64, &back32);
        if (st != MF_OK) return st;

        // every host container is built before the first device allocation, and the plan
        // is held by a guard until it is complete: an exception thrown here (bad_alloc)
        // releases whatever device memory the plan already holds
        std::vector<int32_t> cbHost;
        if (checkBounds && operatorID == MF_OP_ELA)
            cbHost.assign(checkBounds, checkBounds + (size_t)nbNodes * 3);
        std::unique_ptr<mf_plan, void (*)(mf_plan *)> guard(new mf_plan, free_plan);
        mf_plan *p = guard.get();
        p->op = operatorID;
        p->nE = nbElem;
        p->N = nbNodes;
        p->nnz = nnz;
        p->dim = operatorID == MF_OP_ELA ? 9 : 1;
        p->ntiles = (int)h.tiles.size();
        p->block = r.block;
        p->lds_fp64 = lds64;
        p->lds_mixed = lds32;
        p->back_fp64 = back64;
        p->back_mixed = back32;
        p->n_isolated = (int)h.isolated.size();
        p->float_stage = r.float_stage;
        p->max_node = h.max_node;
        p->max_elem = h.max_elem;
        int rc = upload(&p->tiles, h.tiles);
        if (rc == MF_OK) rc = upload(&p->tile_elems, h.tile_elems);
        if (rc == MF_OK) rc = upload_as(&p->tile_verts, h.tile_verts);
        if (rc == MF_OK) rc = upload(&p->tile_nodes, h.tile_nodes);
        if (rc == MF_OK) rc = upload_as(&p->tile_rowinfo, h.tile_rowinfo);
        if (rc == MF_OK) rc = upload(&p->inc, h.inc);
        if (rc == MF_OK) rc = upload_as(&p->task_rec, h.task_rec);
        if (rc == MF_OK) rc = upload(&p->isolated, h.isolated);
        if (rc == MF_OK && !cbHost.empty()) rc = upload(&p->cb, cbHost);
        if (rc == MF_OK)
            rc = operatorID == MF_OP_ELA ? prepare_all<MF_OP_ELA>(p) : prepare_all<MF_OP_LAP>(p);
        if (rc != MF_OK) return rc;  // the guard frees the plan
        fill_stats(h, r.float_stage ? lds32 : lds64, r.block, p->stats);
        *plan = guard.release();
        return MF_OK;
    });
}

extern "C" int mf_plan_host_stats(int operatorID, int nbElem, int nbNodes, int nnz,
                                  const int *elemToNode, const int *row, const int *col,
                                  const double *coord, const mf_plan_params *params,
                                  int64_t out[8])
{
    return mf_guarded([&]() -> int {
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
        fill_stats(h, r.float_stage ? lds32 : lds64, r.block, out);
        return MF_OK;
    });
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

extern "C" int mf_plan_row_order(const mf_plan *plan, int *order, void *stream)
{
    if (!plan || (plan->N > 0 && !order)) return MF_ERR_ARG;
    if (plan->N == 0) return MF_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int grid = std::max(1, std::min(plan->ntiles, 4096));
    hipLaunchKernelGGL(k_row_order, dim3(grid), dim3(256), 0, s, plan->tiles, plan->ntiles,
                       plan->tile_nodes, plan->isolated, plan->n_isolated,
                       plan->N - plan->n_isolated, order);
    MF_CHECK_LAUNCH();
    return MF_OK;
}

extern "C" int mf_plan_write_probe(const mf_plan *plan, double *values, int mode, void *stream)
{
    if (!plan || !values || mode < 0 || mode > 3) return MF_ERR_ARG;
    if (plan->ntiles == 0) return MF_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return MF_ERR_HIP;
    const int per = (plan->ntiles + kXcd - 1) / kXcd;
    const int G = std::max(1, std::min(per, cus / kXcd));  // one workgroup per CU
    const int64_t total = (int64_t)plan->nnz * plan->dim;
    if (plan->dim == 9)
        hipLaunchKernelGGL(k_write_probe<9>, dim3(G * kXcd), dim3(1024), 0, s, plan->tiles,
                           plan->ntiles, G, plan->tile_rowinfo, values, mode, total);
    else
        hipLaunchKernelGGL(k_write_probe<1>, dim3(G * kXcd), dim3(1024), 0, s, plan->tiles,
                           plan->ntiles, G, plan->tile_rowinfo, values, mode, total);
    MF_CHECK_LAUNCH();
    return MF_OK;
}

extern "C" int mf_plan_assemble(mf_plan *plan, const double *coord, const double *coef,
                                double *values, double *prec, int precision, int *singular,
                                void *stream)
{
    if (!plan || !values || !prec || (plan->N > 0 && !coord)) return MF_ERR_ARG;
    if (precision == MF_PREC_FP64 && plan->float_stage) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bool lap = plan->op == MF_OP_LAP;
    switch (precision) {
    case MF_PREC_FP64:
        return lap ? launch_variant<MF_OP_LAP, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s)
                   : launch_variant<MF_OP_ELA, MF_PREC_FP64>(plan, coord, coef, values, prec, singular, s);
    case MF_PREC_MIXED:
        return lap ? launch_variant<MF_OP_LAP, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s)
                   : launch_variant<MF_OP_ELA, MF_PREC_MIXED>(plan, coord, coef, values, prec, singular, s);
    case MF_PREC_MIXED_F32COEF:
        return lap ? launch_variant<MF_OP_LAP, MF_PREC_MIXED_F32COEF>(plan, coord, coef, values, prec, singular, s)
                   : launch_variant<MF_OP_ELA, MF_PREC_MIXED_F32COEF>(plan, coord, coef, values, prec, singular, s);
    default:
        return MF_ERR_ARG;
    }
}

namespace {

// LDS of a launch over several plans: the fixed front sized by their largest tiles with
// ela's coefficient region (k_tiled_jobs keeps one front layout for both operators),
// then the largest back part of any of their tiles.
int jobs_lds(int f_node, int f_elem, int back)
{
    const int verts = mf::align16(f_node * 24);
    const int coef = mf::align16(verts + (f_elem + 1) * 4);
    return coef + mf::align16(f_elem * 16) + back + 16;
}

// Persistent grid of k_tiled_jobs for `lds` bytes of LDS per workgroup: as many
// workgroups per XCD as fit at once, at most one per slot of an XCD's tile range.
// Occupancy is asked once per kernel and LDS size (not a stream operation).
std::mutex g_jobs_mu;
std::map<std::pair<const void *, int>, int> g_jobs_per_cu;
int g_jobs_cus = 0;

template <int PREC, int BT>
int jobs_grid(int lds, int slots, int *G)
{
    const void *kern = reinterpret_cast<const void *>(k_tiled_jobs<PREC, BT>);
    std::lock_guard<std::mutex> lock(g_jobs_mu);
    auto it = g_jobs_per_cu.find({kern, lds});
    if (it == g_jobs_per_cu.end()) {
        int dev = 0, per_cu = 0;
        if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                mf::kMaxLdsBytes) != hipSuccess ||
            hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&g_jobs_cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BT, lds) != hipSuccess)
            return MF_ERR_HIP;
        it = g_jobs_per_cu.emplace(std::make_pair(kern, lds), std::max(1, per_cu)).first;
    }
    *G = std::max(1, std::min(slots, it->second * std::max(1, g_jobs_cus / kXcd)));
    return MF_OK;
}

struct JobsLaunch {
    JobTable table{};
    int ntiles = 0, slots = 0, f_node = 0, f_elem = 0, back = 0, block = 0, lds = 0;
};

int resolve_jobs(const mf_plan_job *jobs, int njobs, int precision, JobsLaunch &L)
{
    if (!jobs || njobs < 1 || njobs > kMaxJobs) return MF_ERR_ARG;
    if (precision != MF_PREC_FP64 && precision != MF_PREC_MIXED &&
        precision != MF_PREC_MIXED_F32COEF)
        return MF_ERR_ARG;
    for (int i = 0; i < njobs; i++) {
        const mf_plan_job &jb = jobs[i];
        const mf_plan *p = jb.plan;
        if (!p || !jb.values || !jb.prec || (p->N > 0 && !jb.coord)) return MF_ERR_ARG;
        if (precision == MF_PREC_FP64 && p->float_stage) return MF_ERR_ARG;
        if (L.block == 0) L.block = p->block;
        if (p->block != L.block) return MF_ERR_ARG;  // one workgroup shape per launch
        L.table.job[i] = make_job(p, jb.coord, jb.coef, jb.values, jb.prec, jb.singular);
        L.ntiles += p->ntiles;
        L.slots += (p->ntiles + kXcd - 1) / kXcd;  // the longest XCD walk holds this many
        if (p->ntiles > 0) {
            L.f_node = std::max(L.f_node, p->max_node);
            L.f_elem = std::max(L.f_elem, p->max_elem);
            L.back = std::max(L.back, precision == MF_PREC_FP64 ? p->back_fp64 : p->back_mixed);
        }
    }
    L.table.njobs = njobs;
    L.lds = L.ntiles > 0 ? jobs_lds(L.f_node, L.f_elem, L.back) : 0;
    return L.lds > mf::kMaxLdsBytes ? MF_ERR_ARG : MF_OK;
}

template <int PREC, int BT>
int launch_jobs(const JobsLaunch &L, hipStream_t s)
{
    int G = 0;
    const int st = jobs_grid<PREC, BT>(L.lds, L.slots, &G);
    if (st != MF_OK) return st;
    auto kern = k_tiled_jobs<PREC, BT>;
    hipLaunchKernelGGL(kern, dim3(G * kXcd), dim3(BT), L.lds, s, L.table, G, L.f_node, L.f_elem);
    MF_CHECK_LAUNCH();
    return MF_OK;
}

template <int PREC>
int launch_jobs_variant(const JobsLaunch &L, hipStream_t s)
{
    switch (L.block) {
    case 512: return launch_jobs<PREC, 512>(L, s);
    case 1024: return launch_jobs<PREC, 1024>(L, s);
    default: return MF_ERR_ARG;
    }
}

}  // namespace

extern "C" int mf_plans_lds_bytes(const mf_plan_job *jobs, int njobs, int precision)
{
    JobsLaunch L;
    const int st = resolve_jobs(jobs, njobs, precision, L);
    return st != MF_OK ? st : L.lds;
}

extern "C" int mf_plans_assemble(const mf_plan_job *jobs, int njobs, int precision, void *stream)
{
    JobsLaunch L;
    int st = resolve_jobs(jobs, njobs, precision, L);
    if (st != MF_OK) return st;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (L.ntiles > 0) {
        st = precision == MF_PREC_FP64    ? launch_jobs_variant<MF_PREC_FP64>(L, s)
             : precision == MF_PREC_MIXED ? launch_jobs_variant<MF_PREC_MIXED>(L, s)
                                          : launch_jobs_variant<MF_PREC_MIXED_F32COEF>(L, s);
        if (st != MF_OK) return st;
    }
    // nodes without elements: the plans that have them add their own small launch
    for (int i = 0; i < njobs; i++) {
        const mf_plan *p = jobs[i].plan;
        if (p->n_isolated == 0) continue;
        if (p->op == MF_OP_ELA)
            hipLaunchKernelGGL(k_prec_isolated<MF_OP_ELA>, dim3((p->n_isolated + 255) / 256),
                               dim3(256), 0, s, p->isolated, p->n_isolated, p->cb, p->N,
                               jobs[i].prec);
        else
            hipLaunchKernelGGL(k_prec_isolated<MF_OP_LAP>, dim3((p->n_isolated + 255) / 256),
                               dim3(256), 0, s, p->isolated, p->n_isolated, p->cb, p->N,
                               jobs[i].prec);
        MF_CHECK_LAUNCH();
    }
    return MF_OK;
}
