// Consumer of the iteration's outputs: block-Jacobi preconditioned conjugate gradients
// on the assembled operator. The reference assembles values and prec every iteration
// but stops before the solve ("solver & update", src/main.cc:353); this closes the loop
// on the GPU.
//
// Operator: the dofs masked by checkBounds (the Dirichlet dofs ela_invert_prec masks,
// src/Fortran/elasclpr.f:29-32; lap masks a node when any of its three flags is set)
// are fixed: K = P A P + (I - P), with P zeroing masked dofs. prec is the block inverse
// of K's diagonal blocks exactly as prec_inversion leaves it (masked rows/columns unit),
// so z = prec r is the block-Jacobi step. Diagonal blocks of ela are symmetric, so the
// Fortran column-major and C row-major views of prec coincide.
//
// One iteration is three streaming launches, each reading what it needs once, and two
// one-workgroup sums of their partials; the scalars stay on the device:
//   q = K p with the partial sums of p.q          (k_spmv, k_spmv_ela_staged: 16 lanes
//                                                   per row; k_sum_pq)
//   x += alpha p, r -= alpha q, z = M r with the partial sums of r.z and r.r
//                                                  (k_update_prec_dot; k_sum_rz)
//   p = z + beta p and its masked copy P p         (k_xpby)
// The product reads each row's blocks as one contiguous stream: the 16 lanes of a row
// take its blocks in turn (consecutive lanes, consecutive 72- or 8-byte blocks), then
// sum their three (one) partial rows in a fixed butterfly; the columns are gathered
// from the masked copy of the direction, one gather per block. Dot products are
// deterministic: a fixed grid, a fixed tree per workgroup, the workgroup partials summed
// in a fixed tree by one workgroup, and only that scalar goes to the host.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "device_common.h"
#include "minifem.h"

namespace {

constexpr int kBT = 256;          // threads per workgroup
constexpr int kGroup = 16;        // lanes per row in the operator product (ela)
// lanes per row of the lap product: 8 took 82 against 87 us per lap 100^3 iteration in the
// plan's order but 217 against 204 us in the shuffled one, 4 was slower in both
constexpr int kLapGroup = 16;
constexpr int kMaxGrid = 2048;    // workgroups of the reducing kernels (partials per sum)
constexpr int kBatch = 8;         // iterations enqueued between two looks at the state

// The iteration's scalars, on the device: the host enqueues iterations without waiting
// for them and reads this every kBatch iterations. Once done is set, every kernel of
// the iterations already enqueued returns at once.
struct CgState {
    double rz, rz1, rr, pq, beta, tol;
    int it, max_it;
    int done;  // 0 running; 1 converged or max_it reached; 2 p.q <= 0 (breakdown)
    int pad;
};

// Per node: bit e set when dof e is fixed (ela); lap: all three bits when any flag is set.
// With an order, entry t describes node order[t].
__global__ void k_mask_bits(const int *__restrict__ cb, int N, bool ela, const int *__restrict__ order,
                            uint8_t *__restrict__ m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int i = order ? order[t] : t;
    uint8_t b = 0;
    if (cb) {
        const int f0 = cb[i] != 0, f1 = cb[(int64_t)N + i] != 0, f2 = cb[2 * (int64_t)N + i] != 0;
        b = ela ? (uint8_t)(f0 | f1 << 1 | f2 << 2) : (uint8_t)((f0 | f1 | f2) ? 7 : 0);
    }
    m[t] = b;
}

// Fixed-order sum of the kBT values of a workgroup in s[]; returns it on thread 0.
__device__ __forceinline__ double block_sum(double *s, double v)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// q = K p over rows i = group, group + groups, ...; part[blockIdx.x] = this workgroup's
// share of p.q. The columns are gathered from pm = P p (fixed dofs already zero), so a
// block costs its own loads and one gather of pm; a fixed row takes p's own value. Every
// lane of a row group runs the same trip count, so the butterfly always has its 16 lanes.
template <bool ELA, int G>
__global__ void __launch_bounds__(kBT)
    k_spmv(const int *__restrict__ row, const int *__restrict__ col,
           const double *__restrict__ values, const uint8_t *__restrict__ mask, int N,
           const double *__restrict__ pm, const double *__restrict__ p, double *__restrict__ q,
           double *__restrict__ part, const CgState *__restrict__ cs)
{
    if (cs && cs->done) return;  // the whole grid returns: nothing waits on the butterfly
    __shared__ double s[kBT];
    constexpr int D = ELA ? 3 : 1;
    const int l = threadIdx.x % G;
    const int groups = gridDim.x * (kBT / G);
    double pq = 0.0;
    for (int i = blockIdx.x * (kBT / G) + threadIdx.x / G; i < N; i += groups) {
        const int q0 = row[i], q1 = row[i + 1];
        const unsigned mi = mask[i];
        double pi[D], acc[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            pi[d] = p[(int64_t)i * D + d];
            acc[d] = 0.0;
        }
        for (int k = q0 + l; k < q1; k += G) {
            const int j = col[k] - 1;
            const double *b = values + (int64_t)k * D * D;
            if constexpr (ELA) {
                double v[9];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const double2 w = *reinterpret_cast<const double2 *>(b + 2 * c);
                    v[2 * c] = w.x;
                    v[2 * c + 1] = w.y;
                }
                v[8] = b[8];
                const double x0 = pm[(int64_t)j * 3], x1 = pm[(int64_t)j * 3 + 1],
                             x2 = pm[(int64_t)j * 3 + 2];
#pragma unroll
                for (int d = 0; d < 3; d++) acc[d] += v[3 * d] * x0 + v[3 * d + 1] * x1 + v[3 * d + 2] * x2;
            } else {
                acc[0] += b[0] * pm[j];
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int d = 0; d < D; d++) acc[d] += __shfl_xor(acc[d], o, G);
        if (l == 0) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double qi = (mi >> d) & 1u ? pi[d] : acc[d];
                q[(int64_t)i * D + d] = qi;
                pq += pi[d] * qi;
            }
        }
    }
    const double t = block_sum(s, pq);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The ela product with the value stream staged through LDS: a wave's four row groups own
// four consecutive rows, whose blocks are one contiguous segment of values; the wave
// reads it in chunks of kChunk blocks with consecutive lanes on consecutive 16-byte
// pieces (the per-lane 72-byte blocks of k_spmv touch each line from several lanes and
// several instructions), then every lane takes its blocks from LDS in k_spmv's order, so
// the sums, and q, are k_spmv's bit for bit. Only this wave reads its LDS region, and LDS
// operations of one wave complete in order, so no workgroup barrier is needed.
constexpr int kChunk = 64;
__global__ void __launch_bounds__(kBT)
    k_spmv_ela_staged(const int *__restrict__ row, const int *__restrict__ col,
                      const double *__restrict__ values, const uint8_t *__restrict__ mask, int N,
                      const double *__restrict__ pm, const double *__restrict__ p,
                      double *__restrict__ q, double *__restrict__ part, const CgState *__restrict__ cs)
{
    if (cs && cs->done) return;
    __shared__ double s[kBT];
    __shared__ __align__(16) double stage[kBT / 64][kChunk * 9];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l = lane & (kGroup - 1), g = lane / kGroup;
    const int stride = gridDim.x * (kBT / kGroup);
    double *st = stage[w];
    double pq = 0.0;
    for (int r0 = blockIdx.x * (kBT / kGroup) + w * (64 / kGroup); r0 < N; r0 += stride) {
        const int i = r0 + g;
        const bool on = i < N;
        const int ic = on ? i : r0;
        const int q0 = row[ic], q1 = on ? row[ic + 1] : q0;
        const int s0 = row[r0], s1 = row[r0 + (64 / kGroup) < N ? r0 + (64 / kGroup) : N];
        const unsigned mi = mask[ic];
        double pi[3], acc[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            pi[d] = p[(int64_t)ic * 3 + d];
            acc[d] = 0.0;
        }
        int k = q0 + l;  // this lane's next block: q0 + l, q0 + l + 16, ...
        for (int c0 = s0; c0 < s1; c0 += kChunk) {
            const int nb = s1 - c0 < kChunk ? s1 - c0 : kChunk, nd = nb * 9;
            const double *src = values + (int64_t)c0 * 9;
            for (int d = 2 * lane; d < nd; d += 128) {
                if (d + 1 < nd) {
                    const double2 v = *reinterpret_cast<const double2 *>(src + d);
                    st[d] = v.x;
                    st[d + 1] = v.y;
                } else {
                    st[d] = src[d];
                }
            }
            __builtin_amdgcn_wave_barrier();
            for (; k < q1 && k < c0 + nb; k += kGroup) {
                const int j = col[k] - 1;
                const double *b = st + (k - c0) * 9;
                const double x0 = pm[(int64_t)j * 3], x1 = pm[(int64_t)j * 3 + 1],
                             x2 = pm[(int64_t)j * 3 + 2];
#pragma unroll
                for (int d = 0; d < 3; d++)
                    acc[d] += b[3 * d] * x0 + b[3 * d + 1] * x1 + b[3 * d + 2] * x2;
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int o = kGroup / 2; o > 0; o >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) acc[d] += __shfl_xor(acc[d], o, kGroup);
        if (l == 0 && on) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double qi = (mi >> d) & 1u ? pi[d] : acc[d];
                q[(int64_t)i * 3 + d] = qi;
                pq += pi[d] * qi;
            }
        }
    }
    const double t = block_sum(s, pq);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// z = M r per node block; partA/partB[blockIdx.x] = this workgroup's share of r.z, r.r
template <bool ELA>
__global__ void __launch_bounds__(kBT)
    k_prec_dot(const double *__restrict__ prec, const uint8_t *__restrict__ mask, int N,
               const double *__restrict__ r, double *__restrict__ z, double *__restrict__ partA,
               double *__restrict__ partB)
{
    __shared__ double s[kBT];
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * kBT + threadIdx.x; i < N; i += gridDim.x * kBT) {
        if constexpr (ELA) {
            const double *m = prec + (int64_t)i * 9;
            const double r0 = r[(int64_t)i * 3], r1 = r[(int64_t)i * 3 + 1], r2 = r[(int64_t)i * 3 + 2];
            const double z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
            const double z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
            const double z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
            z[(int64_t)i * 3] = z0;
            z[(int64_t)i * 3 + 1] = z1;
            z[(int64_t)i * 3 + 2] = z2;
            rz += r0 * z0 + r1 * z1 + r2 * z2;
            rr += r0 * r0 + r1 * r1 + r2 * r2;
        } else {
            // a masked lap node's row of K is the identity, so its prec is 1
            const double ri = r[i], zi = mask[i] ? ri : prec[i] * ri;
            z[i] = zi;
            rz += ri * zi;
            rr += ri * ri;
        }
    }
    const double a = block_sum(s, rz);
    const double b = block_sum(s, rr);
    if (threadIdx.x == 0) {
        partA[blockIdx.x] = a;
        partB[blockIdx.x] = b;
    }
}

// out[k] = the fixed-order sum of part_k[0..n) (k < npart), one workgroup
__global__ void __launch_bounds__(kBT)
    k_sum_parts(const double *__restrict__ partA, const double *__restrict__ partB, int n,
                double *__restrict__ out)
{
    __shared__ double s[kBT];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += kBT) {
        a += partA[i];
        if (partB) b += partB[i];
    }
    const double ta = block_sum(s, a);
    const double tb = partB ? block_sum(s, b) : 0.0;
    if (threadIdx.x == 0) {
        out[0] = ta;
        out[1] = tb;
    }
}

// x += alpha p, r -= alpha q (alpha = r.z / p.q), then z = M r with the partial sums of
// r.z and r.r: k_prec_dot on the updated residual, in the same grid and order
template <bool ELA>
__global__ void __launch_bounds__(kBT)
    k_update_prec_dot(const CgState *__restrict__ cs, const double *__restrict__ p,
                      const double *__restrict__ q, double *__restrict__ x, double *__restrict__ r,
                      const double *__restrict__ prec, const uint8_t *__restrict__ mask, int N,
                      double *__restrict__ z, double *__restrict__ partA, double *__restrict__ partB)
{
    if (cs->done) return;
    __shared__ double s[kBT];
    const double alpha = cs->rz / cs->pq;
    double rz = 0.0, rr = 0.0;
    for (int i = blockIdx.x * kBT + threadIdx.x; i < N; i += gridDim.x * kBT) {
        if constexpr (ELA) {
            double rv[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int64_t k = (int64_t)i * 3 + d;
                x[k] += alpha * p[k];
                rv[d] = r[k] - alpha * q[k];
                r[k] = rv[d];
            }
            const double *m = prec + (int64_t)i * 9;
            const double z0 = m[0] * rv[0] + m[1] * rv[1] + m[2] * rv[2];
            const double z1 = m[3] * rv[0] + m[4] * rv[1] + m[5] * rv[2];
            const double z2 = m[6] * rv[0] + m[7] * rv[1] + m[8] * rv[2];
            z[(int64_t)i * 3] = z0;
            z[(int64_t)i * 3 + 1] = z1;
            z[(int64_t)i * 3 + 2] = z2;
            rz += rv[0] * z0 + rv[1] * z1 + rv[2] * z2;
            rr += rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        } else {
            x[i] += alpha * p[i];
            const double ri = r[i] - alpha * q[i];
            r[i] = ri;
            const double zi = mask[i] ? ri : prec[i] * ri;
            z[i] = zi;
            rz += ri * zi;
            rr += ri * ri;
        }
    }
    const double a = block_sum(s, rz);
    const double b = block_sum(s, rr);
    if (threadIdx.x == 0) {
        partA[blockIdx.x] = a;
        partB[blockIdx.x] = b;
    }
}

// the iteration's p.q from the spmv partials; p.q <= 0 stops the iteration before x moves
__global__ void __launch_bounds__(kBT) k_sum_pq(const double *__restrict__ part, int n, CgState *cs)
{
    if (cs->done) return;
    __shared__ double s[kBT];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += kBT) a += part[i];
    const double t = block_sum(s, a);
    if (threadIdx.x == 0) {
        cs->pq = t;
        if (!(t > 0.0)) cs->done = 2;  // K is not positive definite on p (or p = 0)
    }
}

// r.z and r.r of the updated residual from the partials, and the end of the iteration:
// beta = r.z (new) / r.z (old) for the direction update, the new r.z becomes the old one,
// and the solve stops at max_it or convergence (the direction update is then not needed)
__global__ void __launch_bounds__(kBT)
    k_sum_rz(const double *__restrict__ partA, const double *__restrict__ partB, int n, CgState *cs)
{
    if (cs->done) return;
    __shared__ double s[kBT];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += kBT) {
        a += partA[i];
        b += partB[i];
    }
    const double ta = block_sum(s, a);
    const double tb = block_sum(s, b);
    if (threadIdx.x == 0) {
        cs->rz1 = ta;
        cs->rr = tb;
        cs->beta = ta / cs->rz;
        cs->rz = ta;
        cs->it += 1;
        if (cs->it >= cs->max_it || !(sqrt(tb) > cs->tol)) cs->done = 1;
    }
}

// p = z + beta p and pm = P p; without a state: p = z, whatever p held before
__global__ void k_xpby(const double *__restrict__ z, const CgState *__restrict__ cs,
                       const uint8_t *__restrict__ mask, int D, double *__restrict__ p,
                       double *__restrict__ pm, int64_t n)
{
    if (cs && cs->done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = cs ? z[i] + cs->beta * p[i] : z[i];
    p[i] = v;
    pm[i] = (mask[i / D] >> (i % D)) & 1u ? 0.0 : v;
}

// rhs = P b (fixed dofs set to 0); r = rhs - q; partA/partB: shares of rhs.rhs
__global__ void __launch_bounds__(kBT)
    k_residual0(const double *__restrict__ b, const uint8_t *__restrict__ mask, int D,
                const double *__restrict__ q, double *__restrict__ rhs, double *__restrict__ r,
                int64_t n, double *__restrict__ part)
{
    __shared__ double s[kBT];
    double bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBT) {
        const int node = (int)(i / D), d = (int)(i % D);
        const double v = (mask[node] >> d) & 1u ? 0.0 : b[i];
        rhs[i] = v;
        r[i] = v - q[i];
        bb += v * v;
    }
    const double t = block_sum(s, bb);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- the system in another node order (mf_pcg_ordered)

// flag = 1 when some order[t] lies outside [0, N)
__global__ void k_order_range(const int *__restrict__ order, int N, int *flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N && (order[t] < 0 || order[t] >= N)) atomicOr(flag, 1);
}

__global__ void k_order_invert(const int *__restrict__ order, int N, int *__restrict__ inv)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) inv[order[t]] = t;
}

// flag = 1 when order repeats a node (then some inv[order[t]] names another t)
__global__ void k_order_dups(const int *__restrict__ order, const int *__restrict__ inv, int N, int *flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N && inv[order[t]] != t) atomicOr(flag, 1);
}

// len[t] = length of row order[t]; len[N] = 0 (the exclusive scan then ends in nnz);
// flag = 1 for a negative length
__global__ void k_order_rowlen(const int *__restrict__ order, const int *__restrict__ row, int N,
                               int *__restrict__ len, int *flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < N) {
        const int l = row[order[t] + 1] - row[order[t]];
        len[t] = l;
        if (l < 0) atomicOr(flag, 1);
    } else if (t == N) {
        len[t] = 0;
    }
}

// flag = 1 for a column outside 1..N
__global__ void k_check_cols(const int *__restrict__ col, int64_t nnz, int N, int *flag)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz && (col[k] < 1 || col[k] > N)) atomicOr(flag, 1);
}

// new row t = old row order[t], its entries in their own order, columns renumbered
template <int DD>
__global__ void __launch_bounds__(kBT)
    k_order_csr(const int *__restrict__ order, const int *__restrict__ inv,
                const int *__restrict__ row, const int *__restrict__ col,
                const double *__restrict__ values, const int *__restrict__ rowp, int N,
                int *__restrict__ colp, double *__restrict__ valp)
{
    const int l = threadIdx.x % kGroup;
    const int groups = gridDim.x * (kBT / kGroup);
    for (int t = blockIdx.x * (kBT / kGroup) + threadIdx.x / kGroup; t < N; t += groups) {
        const int i = order[t], src = row[i], len = row[i + 1] - src, dst = rowp[t];
        for (int k = l; k < len; k += kGroup) {
            colp[dst + k] = inv[col[src + k] - 1] + 1;
#pragma unroll
            for (int w = 0; w < DD; w++)
                valp[(int64_t)(dst + k) * DD + w] = values[(int64_t)(src + k) * DD + w];
        }
    }
}

// out[t W + w] = in[order[t] W + w] (gather) or the reverse (scatter)
__global__ void k_order_rows(const int *__restrict__ order, int64_t total, int W, bool gather,
                             const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const int64_t t = k / W, w = k % W, o = (int64_t)order[t] * W + w;
    if (gather) out[k] = in[o];
    else out[o] = in[k];
}

inline dim3 g1(int64_t n) { return dim3((unsigned)((n + kBT - 1) / kBT)); }
inline int grid_for(int64_t items) { return (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, (items + kBT - 1) / kBT)); }

}  // namespace

// The solve on one numbering of the nodes; order (may be NULL) maps it to checkBounds'
// (entry t of every array is node order[t]).
static int pcg_core(bool ela, int nbNodes, const int *row, const int *col, const double *values,
             const double *prec, const int *checkBounds, const int *order, const double *b,
             double *x, double rtol, int maxIter, int *iterations, double *relResidual,
             hipStream_t s)
{
    const int D = ela ? 3 : 1;
    const int64_t n = (int64_t)nbNodes * D;
    if (iterations) *iterations = 0;
    if (relResidual) *relResidual = 0.0;
    if (n == 0) return MF_OK;
    // r, z, p, q, rhs, pm = P p; partials (2 x kMaxGrid); the two sums; the state; the
    // mask bytes
    const size_t vec = (size_t)n * sizeof(double);
    const size_t bytes = 6 * vec + 2 * kMaxGrid * sizeof(double) + 2 * sizeof(double) +
                         sizeof(CgState) + (size_t)nbNodes;
    unsigned char *buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return MF_ERR_NOMEM;
    double *r = reinterpret_cast<double *>(buf), *z = r + n, *p = z + n, *q = p + n, *rhs = q + n;
    double *pm = rhs + n, *partA = pm + n, *partB = partA + kMaxGrid, *sums = partB + kMaxGrid;
    CgState *cs = reinterpret_cast<CgState *>(sums + 2);
    uint8_t *mask = reinterpret_cast<uint8_t *>(cs + 1);
    const int gSpmv = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, ((int64_t)nbNodes * kGroup + kBT - 1) / kBT));
    const int gSpmvLap = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, ((int64_t)nbNodes * kLapGroup + kBT - 1) / kBT));
    const int gNode = grid_for(nbNodes), gDof = grid_for(n);
    int st = MF_OK;
    auto launched = [&]() {
        if (st == MF_OK && hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
        return st == MF_OK;
    };
    // out = K in, from in and its masked copy in_m (state: skipped once the solve is done)
    auto spmv = [&](const double *in_m, const double *in, double *out, const CgState *c) {
        if (ela)
            hipLaunchKernelGGL(k_spmv_ela_staged, dim3(gSpmv), dim3(kBT), 0, s, row, col, values,
                               mask, nbNodes, in_m, in, out, partA, c);
        else
            hipLaunchKernelGGL((k_spmv<false, kLapGroup>), dim3(gSpmvLap), dim3(kBT), 0, s, row, col,
                               values, mask, nbNodes, in_m, in, out, partA, c);
    };
    // p = z (c NULL) or p = z + beta p, and pm = P p
    auto xpby = [&](const CgState *c) {
        hipLaunchKernelGGL(k_xpby, g1(n), dim3(kBT), 0, s, z, c, mask, D, p, pm, n);
    };
    auto prec_dot = [&]() {
        if (ela)
            hipLaunchKernelGGL(k_prec_dot<true>, dim3(gNode), dim3(kBT), 0, s, prec, mask, nbNodes, r,
                               z, partA, partB);
        else
            hipLaunchKernelGGL(k_prec_dot<false>, dim3(gNode), dim3(kBT), 0, s, prec, mask, nbNodes,
                               r, z, partA, partB);
    };
    // the two sums of the partials of a grid of g workgroups, on the host
    double h2[2] = {0.0, 0.0};
    auto sums_of = [&](int g, bool two) {
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBT), 0, s, partA, two ? partB : nullptr, g,
                           sums);
        if (!launched()) return false;
        if (hipMemcpyAsync(h2, sums, sizeof(h2), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = MF_ERR_HIP;
        return st == MF_OK;
    };

    // rhs = P b; r = rhs - K x; z = M r; p = z
    hipLaunchKernelGGL(k_mask_bits, g1(nbNodes), dim3(kBT), 0, s, checkBounds, nbNodes, ela, order,
                       mask);
    launched();
    double bnorm2 = 0.0, rz = 0.0, rr = 0.0;
    if (st == MF_OK) {
        // p = x, pm = P x: scratch for q = K x
        hipLaunchKernelGGL(k_xpby, g1(n), dim3(kBT), 0, s, x, nullptr, mask, D, p, pm, n);
        spmv(pm, p, q, nullptr);  // its p.q partials are not used
        hipLaunchKernelGGL(k_residual0, dim3(gDof), dim3(kBT), 0, s, b, mask, D, q, rhs, r, n,
                           partA);
        if (launched() && sums_of(gDof, false)) bnorm2 = h2[0];
    }
    if (st == MF_OK) {
        prec_dot();
        if (launched() && sums_of(gNode, true)) {
            rz = h2[0];
            rr = h2[1];
        }
    }
    if (st == MF_OK) {
        xpby(nullptr);  // p = z
        launched();
    }
    const double bnorm = std::sqrt(bnorm2) > 0.0 ? std::sqrt(bnorm2) : 1.0;
    // the iterations, enqueued kBatch at a time; k_sum_rz counts each one and decides, as
    // the loop condition would, whether another one runs
    CgState h{};
    h.rz = rz;
    h.rz1 = rz;
    h.rr = rr;
    h.tol = rtol * bnorm;
    h.it = 0;
    h.max_it = maxIter;
    h.done = maxIter > 0 && std::sqrt(rr) > h.tol ? 0 : 1;
    if (st == MF_OK && (hipMemcpyAsync(cs, &h, sizeof(h), hipMemcpyHostToDevice, s) != hipSuccess ||
                        hipStreamSynchronize(s) != hipSuccess))
        st = MF_ERR_HIP;
    while (st == MF_OK && !h.done) {
        for (int k = 0; k < kBatch; k++) {
            spmv(pm, p, q, cs);
            hipLaunchKernelGGL(k_sum_pq, dim3(1), dim3(kBT), 0, s, partA, ela ? gSpmv : gSpmvLap, cs);
            if (ela)
                hipLaunchKernelGGL(k_update_prec_dot<true>, dim3(gNode), dim3(kBT), 0, s, cs, p, q,
                                   x, r, prec, mask, nbNodes, z, partA, partB);
            else
                hipLaunchKernelGGL(k_update_prec_dot<false>, dim3(gNode), dim3(kBT), 0, s, cs, p, q,
                                   x, r, prec, mask, nbNodes, z, partA, partB);
            hipLaunchKernelGGL(k_sum_rz, dim3(1), dim3(kBT), 0, s, partA, partB, gNode, cs);
            xpby(cs);
        }
        if (!launched()) break;
        if (hipMemcpyAsync(&h, cs, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = MF_ERR_HIP;
    }
    const int it = h.it;
    rr = h.rr;
    if (st == MF_OK && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    if (iterations) *iterations = it;
    if (relResidual) *relResidual = std::sqrt(rr) / bnorm;
    (void)hipStreamSynchronize(s);  // nothing in flight reads buf when it is freed
    (void)hipFree(buf);
    return st;
}

extern "C" int mf_pcg(int operatorID, int nbNodes, const int *row, const int *col,
                      const double *values, const double *prec, const int *checkBounds,
                      const double *b, double *x, double rtol, int maxIter, int *iterations,
                      double *relResidual, void *stream)
{
    const bool ela = operatorID == MF_OP_ELA;
    if ((operatorID != MF_OP_LAP && !ela) || nbNodes < 0 || !row || !values || !prec || !b || !x ||
        maxIter < 0 || !(rtol > 0.0))
        return MF_ERR_ARG;
    return pcg_core(ela, nbNodes, row, col, values, prec, checkBounds, nullptr, b, x, rtol, maxIter,
                    iterations, relResidual, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int mf_pcg_ordered(int operatorID, int nbNodes, const int *row, const int *col,
                              const double *values, const double *prec, const int *checkBounds,
                              const int *order, const double *b, double *x, double rtol,
                              int maxIter, int *iterations, double *relResidual, void *stream)
{
    const bool ela = operatorID == MF_OP_ELA;
    if ((operatorID != MF_OP_LAP && !ela) || nbNodes < 0 || !row || !values || !prec || !b || !x ||
        !order || maxIter < 0 || !(rtol > 0.0))
        return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (iterations) *iterations = 0;
    if (relResidual) *relResidual = 0.0;
    const int N = nbNodes, D = ela ? 3 : 1, DD = D * D;
    if (N == 0) return MF_OK;
    int ends[2] = {0, 0};  // row[0], row[N]
    if (hipMemcpyAsync(&ends[0], row, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&ends[1], row + N, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MF_ERR_HIP;
    const int nnz = ends[1];
    if (ends[0] != 0 || nnz < 0 || (nnz > 0 && !col)) return MF_ERR_ARG;
    size_t scan_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int *)nullptr, (int *)nullptr,
                                         N + 1, s) != hipSuccess)
        return MF_ERR_HIP;
    // flag, inv[N], len[N+1], rowp[N+1], colp[nnz], valp[nnz DD], precp[N DD], bp, xp[N D],
    // the scan's scratch; each piece 256-byte aligned
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t o_inv = 256, o_len = o_inv + al(4 * (size_t)N), o_rowp = o_len + al(4 * ((size_t)N + 1)),
                 o_col = o_rowp + al(4 * ((size_t)N + 1)), o_val = o_col + al(4 * (size_t)nnz),
                 o_prec = o_val + al(8 * (size_t)nnz * DD), o_b = o_prec + al(8 * (size_t)N * DD),
                 o_x = o_b + al(8 * (size_t)N * D), o_scan = o_x + al(8 * (size_t)N * D),
                 bytes = o_scan + al(scan_bytes);
    unsigned char *buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return MF_ERR_NOMEM;
    int *flag = reinterpret_cast<int *>(buf), *inv = reinterpret_cast<int *>(buf + o_inv),
        *len = reinterpret_cast<int *>(buf + o_len), *rowp = reinterpret_cast<int *>(buf + o_rowp),
        *colp = reinterpret_cast<int *>(buf + o_col);
    double *valp = reinterpret_cast<double *>(buf + o_val), *precp = reinterpret_cast<double *>(buf + o_prec),
           *bp = reinterpret_cast<double *>(buf + o_b), *xp = reinterpret_cast<double *>(buf + o_x);
    int st = MF_OK;
    auto launched = [&]() {
        if (st == MF_OK && hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
        return st == MF_OK;
    };
    // the copies below write where row and order say: order must be a permutation of
    // 0..N-1 (every index in range first, then no repeats), every row length >= 0 (with
    // row[0] = 0 and row[N] = nnz, each new row then lies inside 0..nnz) and every column
    // in 1..N
    auto flag_set = [&]() {
        int h = 0;
        if (hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            st = MF_ERR_HIP;
            return true;
        }
        return h != 0;
    };
    const int zero = 0;
    if (hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_OK) {
        hipLaunchKernelGGL(k_order_range, g1(N), dim3(kBT), 0, s, order, N, flag);
        if (launched() && flag_set() && st == MF_OK) st = MF_ERR_ARG;
    }
    if (st == MF_OK) {
        hipLaunchKernelGGL(k_order_invert, g1(N), dim3(kBT), 0, s, order, N, inv);
        hipLaunchKernelGGL(k_order_dups, g1(N), dim3(kBT), 0, s, order, inv, N, flag);
        if (launched() && flag_set() && st == MF_OK) st = MF_ERR_ARG;
    }
    if (st == MF_OK) {
        hipLaunchKernelGGL(k_order_rowlen, g1((int64_t)N + 1), dim3(kBT), 0, s, order, row, N, len,
                           flag);
        if (nnz > 0)
            hipLaunchKernelGGL(k_check_cols, g1(nnz), dim3(kBT), 0, s, col, (int64_t)nnz, N, flag);
        if (launched() && flag_set() && st == MF_OK) st = MF_ERR_ARG;
    }
    if (st == MF_OK && hipcub::DeviceScan::ExclusiveSum(buf + o_scan, scan_bytes, len, rowp, N + 1,
                                                        s) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_OK) {
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, ((int64_t)N * kGroup + kBT - 1) / kBT));
        if (ela)
            hipLaunchKernelGGL(k_order_csr<9>, dim3(g), dim3(kBT), 0, s, order, inv, row, col, values,
                               rowp, N, colp, valp);
        else
            hipLaunchKernelGGL(k_order_csr<1>, dim3(g), dim3(kBT), 0, s, order, inv, row, col, values,
                               rowp, N, colp, valp);
        hipLaunchKernelGGL(k_order_rows, g1((int64_t)N * DD), dim3(kBT), 0, s, order, (int64_t)N * DD,
                           DD, true, prec, precp);
        hipLaunchKernelGGL(k_order_rows, g1((int64_t)N * D), dim3(kBT), 0, s, order, (int64_t)N * D, D,
                           true, b, bp);
        hipLaunchKernelGGL(k_order_rows, g1((int64_t)N * D), dim3(kBT), 0, s, order, (int64_t)N * D, D,
                           true, x, xp);
        launched();
    }
    if (st == MF_OK)
        st = pcg_core(ela, N, rowp, colp, valp, precp, checkBounds, order, bp, xp, rtol, maxIter,
                      iterations, relResidual, s);
    if (st == MF_OK) {
        hipLaunchKernelGGL(k_order_rows, g1((int64_t)N * D), dim3(kBT), 0, s, order, (int64_t)N * D, D,
                           false, xp, x);
        if (launched() && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    }
    (void)hipStreamSynchronize(s);  // nothing in flight reads buf when it is freed
    (void)hipFree(buf);
    return st;
}
