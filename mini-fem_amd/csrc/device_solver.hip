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
// Dot products are deterministic: fixed grid, fixed per-block tree, partials summed in
// block order on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "device_common.h"
#include "minifem.h"

namespace {

constexpr int kDotBlocks = 1024, kBT = 256;

// masked dof flag: ela per dof, lap per node (any flag)
__device__ __forceinline__ bool masked(const int *cb, int N, int node, int d, bool ela)
{
    if (!cb) return false;
    if (ela) return cb[(int64_t)d * N + node] != 0;
    return (cb[node] | cb[(int64_t)N + node] | cb[2 * (int64_t)N + node]) != 0;
}

// y = K x, one thread per node
template <bool ELA>
__global__ void k_spmv(const int *__restrict__ row, const int *__restrict__ col,
                       const double *__restrict__ values, const int *__restrict__ cb, int N,
                       const double *__restrict__ x, double *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int D = ELA ? 3 : 1;
    double acc[D];
#pragma unroll
    for (int d = 0; d < D; d++) acc[d] = 0.0;
    for (int q = row[i]; q < row[i + 1]; q++) {
        const int j = col[q] - 1;
        double xj[D];
#pragma unroll
        for (int e = 0; e < D; e++) xj[e] = masked(cb, N, j, e, ELA) ? 0.0 : x[(int64_t)j * D + e];
        const double *b = values + (int64_t)q * D * D;
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
            for (int e = 0; e < D; e++) acc[d] += b[d * D + e] * xj[e];
    }
#pragma unroll
    for (int d = 0; d < D; d++)
        y[(int64_t)i * D + d] = masked(cb, N, i, d, ELA) ? x[(int64_t)i * D + d] : acc[d];
}

// z = prec r (per node block)
template <bool ELA>
__global__ void k_prec_apply(const double *__restrict__ prec, int N, const double *__restrict__ r,
                             double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if constexpr (ELA) {
        const double *p = prec + (int64_t)i * 9;
        const double r0 = r[i * 3], r1 = r[i * 3 + 1], r2 = r[i * 3 + 2];
        z[i * 3 + 0] = p[0] * r0 + p[1] * r1 + p[2] * r2;
        z[i * 3 + 1] = p[3] * r0 + p[4] * r1 + p[5] * r2;
        z[i * 3 + 2] = p[6] * r0 + p[7] * r1 + p[8] * r2;
    } else {
        z[i] = prec[i] * r[i];
    }
}

// lap with a mask: the masked node's row of K is the identity, so its prec is 1
__global__ void k_prec_apply_lap_masked(const double *__restrict__ prec, const int *__restrict__ cb,
                                        int N, const double *__restrict__ r, double *__restrict__ z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    z[i] = masked(cb, N, i, 0, false) ? r[i] : prec[i] * r[i];
}

// per-block partial sums of a . b over a fixed grid (grid-stride, then a tree)
__global__ void __launch_bounds__(kBT)
    k_dot(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
          double *__restrict__ part)
{
    __shared__ double s[kBT];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x; i < n; i += (int64_t)kDotBlocks * kBT)
        acc += a[i] * b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// y += alpha x
__global__ void k_axpy(double alpha, const double *__restrict__ x, double *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += alpha * x[i];
}

// p = z + beta p
__global__ void k_xpby(const double *__restrict__ z, double beta, double *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// b_masked = P b (masked dofs of the right-hand side set to the fixed value 0)
__global__ void k_mask_vec(const int *__restrict__ cb, int N, bool ela, double *__restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int D = ela ? 3 : 1;
    for (int d = 0; d < D; d++)
        if (masked(cb, N, i, d, ela)) v[(int64_t)i * D + d] = 0.0;
}

struct Dot {
    double *part = nullptr;
    std::vector<double> h = std::vector<double>(kDotBlocks);
    int run(const double *a, const double *b, int64_t n, hipStream_t s, double *out)
    {
        hipLaunchKernelGGL(k_dot, dim3(kDotBlocks), dim3(kBT), 0, s, a, b, n, part);
        if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
        if (hipMemcpyAsync(h.data(), part, kDotBlocks * sizeof(double), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return MF_ERR_HIP;
        double acc = 0.0;
        for (double v : h) acc += v;
        *out = acc;
        return MF_OK;
    }
};

inline dim3 g1(int64_t n) { return dim3((unsigned)((n + kBT - 1) / kBT)); }

}  // namespace

extern "C" int mf_pcg(int operatorID, int nbNodes, const int *row, const int *col,
                      const double *values, const double *prec, const int *checkBounds,
                      const double *b, double *x, double rtol, int maxIter, int *iterations,
                      double *relResidual, void *stream)
{
    const bool ela = operatorID == MF_OP_ELA;
    if ((operatorID != MF_OP_LAP && !ela) || nbNodes < 0 || !row || !values || !prec || !b || !x ||
        maxIter < 0 || !(rtol > 0.0))
        return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t n = (int64_t)nbNodes * (ela ? 3 : 1);
    if (iterations) *iterations = 0;
    if (relResidual) *relResidual = 0.0;
    if (n == 0) return MF_OK;
    double *buf = nullptr;  // r, z, p, q, rhs
    if (hipMalloc(&buf, (size_t)n * 5 * sizeof(double)) != hipSuccess) return MF_ERR_NOMEM;
    Dot dot;
    if (hipMalloc(&dot.part, kDotBlocks * sizeof(double)) != hipSuccess) {
        (void)hipFree(buf);
        return MF_ERR_NOMEM;
    }
    double *r = buf, *z = buf + n, *p = buf + 2 * n, *q = buf + 3 * n, *rhs = buf + 4 * n;
    int st = MF_OK;
    auto spmv = [&](const double *in, double *out) {
        if (ela)
            hipLaunchKernelGGL(k_spmv<true>, g1(nbNodes), dim3(kBT), 0, s, row, col, values,
                               checkBounds, nbNodes, in, out);
        else
            hipLaunchKernelGGL(k_spmv<false>, g1(nbNodes), dim3(kBT), 0, s, row, col, values,
                               checkBounds, nbNodes, in, out);
    };
    auto papply = [&](const double *in, double *out) {
        if (ela)
            hipLaunchKernelGGL(k_prec_apply<true>, g1(nbNodes), dim3(kBT), 0, s, prec, nbNodes, in, out);
        else if (checkBounds)
            hipLaunchKernelGGL(k_prec_apply_lap_masked, g1(nbNodes), dim3(kBT), 0, s, prec,
                               checkBounds, nbNodes, in, out);
        else
            hipLaunchKernelGGL(k_prec_apply<false>, g1(nbNodes), dim3(kBT), 0, s, prec, nbNodes, in, out);
    };
    // rhs = P b; r = rhs - K x; z = M r; p = z
    if (hipMemcpyAsync(rhs, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_OK && checkBounds)
        hipLaunchKernelGGL(k_mask_vec, g1(nbNodes), dim3(kBT), 0, s, checkBounds, nbNodes, ela, rhs);
    if (st == MF_OK) {
        spmv(x, q);
        if (hipMemcpyAsync(r, rhs, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = MF_ERR_HIP;
        hipLaunchKernelGGL(k_axpy, g1(n), dim3(kBT), 0, s, -1.0, q, r, n);
        papply(r, z);
        if (hipMemcpyAsync(p, z, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess)
            st = MF_ERR_HIP;
        if (hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
    }
    double bnorm2 = 0.0, rz = 0.0, rr = 0.0;
    if (st == MF_OK) st = dot.run(rhs, rhs, n, s, &bnorm2);
    if (st == MF_OK) st = dot.run(r, z, n, s, &rz);
    if (st == MF_OK) st = dot.run(r, r, n, s, &rr);
    const double bnorm = std::sqrt(bnorm2) > 0.0 ? std::sqrt(bnorm2) : 1.0;
    int it = 0;
    while (st == MF_OK && it < maxIter && std::sqrt(rr) > rtol * bnorm) {
        spmv(p, q);
        double pq = 0.0;
        if ((st = dot.run(p, q, n, s, &pq)) != MF_OK) break;
        if (!(pq > 0.0)) break;  // K is not positive definite on p (or p = 0)
        const double alpha = rz / pq;
        hipLaunchKernelGGL(k_axpy, g1(n), dim3(kBT), 0, s, alpha, p, x, n);
        hipLaunchKernelGGL(k_axpy, g1(n), dim3(kBT), 0, s, -alpha, q, r, n);
        papply(r, z);
        if (hipGetLastError() != hipSuccess) {
            st = MF_ERR_HIP;
            break;
        }
        double rz1 = 0.0;
        if ((st = dot.run(r, z, n, s, &rz1)) != MF_OK) break;
        if ((st = dot.run(r, r, n, s, &rr)) != MF_OK) break;
        hipLaunchKernelGGL(k_xpby, g1(n), dim3(kBT), 0, s, z, rz1 / rz, p, n);
        rz = rz1;
        it++;
    }
    if (st == MF_OK && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    if (iterations) *iterations = it;
    if (relResidual) *relResidual = std::sqrt(rr) / bnorm;
    (void)hipFree(dot.part);
    (void)hipFree(buf);
    return st;
}
