// Device-side element math shared by every assembly schedule, and the 3x3
// block-Jacobi inverse. All arithmetic is written in the reference's evaluation order
// and compiled with -ffp-contract=off, so each product and sum rounds separately, as
// on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define MF_CHECK_LAUNCH()                                       \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return MF_ERR_HIP;                \
    } while (0)

namespace mf {

// Scaled P1 gradients of one tet, src/assembly.cc:85-121: a=v0-v3, b=v2-v3, c=v1-v3,
// g0=b x c, g1=a x b, g2=c x a, g3=-(g0+g1+g2), det=a.g0, g *= (1./det).
__device__ __forceinline__ void elem_grad(const double (&c)[4][3], double (&g)[4][3])
{
    const double xa = c[0][0] - c[3][0], xb = c[2][0] - c[3][0], xc = c[1][0] - c[3][0];
    const double ya = c[0][1] - c[3][1], yb = c[2][1] - c[3][1], yc = c[1][1] - c[3][1];
    const double za = c[0][2] - c[3][2], zb = c[2][2] - c[3][2], zc = c[1][2] - c[3][2];
    g[0][0] = yb * zc - yc * zb;
    g[0][1] = zb * xc - zc * xb;
    g[0][2] = xb * yc - xc * yb;
    g[1][0] = ya * zb - yb * za;
    g[1][1] = za * xb - zb * xa;
    g[1][2] = xa * yb - xb * ya;
    g[2][0] = yc * za - ya * zc;
    g[2][1] = zc * xa - za * xc;
    g[2][2] = xc * ya - xa * yc;
#pragma unroll
    for (int d = 0; d < 3; d++) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
    const double vol = xa * g[0][0] + ya * g[0][1] + za * g[0][2];
    const double r = 1. / vol;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int d = 0; d < 3; d++) g[i][d] *= r;
}

__device__ __forceinline__ void load_coords(const double *__restrict__ coord, int4 v,
                                            double (&c)[4][3])
{
    const int n[4] = {v.x - 1, v.y - 1, v.z - 1, v.w - 1};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double *p = coord + (int64_t)n[i] * 3;
        c[i][0] = p[0];
        c[i][1] = p[1];
        c[i][2] = p[2];
    }
}

// Per-element operator constants: lap k; ela (lambda+2mu, lambda+mu, mu). NULL coef
// gives the reference constants (1 | 2.25, 1.25, 1.0), src/assembly.cc:386-409.
struct ElaCoef {
    double cd, co, cm;
};

__device__ __forceinline__ ElaCoef ela_coef(const double *__restrict__ coef, int e)
{
    ElaCoef k{2.25, 1.25, 1.0};
    if (coef) {
        const double2 lm = reinterpret_cast<const double2 *>(coef)[e];
        k.cd = lm.x + 2.0 * lm.y;
        k.co = lm.x + lm.y;
        k.cm = lm.y;
    }
    return k;
}

// One 3x3 ela block (row-major) of the (j,k) vertex pair, src/assembly.cc:386-409.
template <typename T>
__device__ __forceinline__ void ela_block(const T *gj, const T *gk, T cd, T co, T cm, T (&b)[9])
{
    const T p00 = gj[0] * gk[0], p11 = gj[1] * gk[1], p22 = gj[2] * gk[2];
    b[0] = p00 * cd + p11 * cm + p22 * cm;
    b[1] = gj[0] * gk[1] * co;
    b[2] = gj[0] * gk[2] * co;
    b[3] = gj[1] * gk[0] * co;
    b[4] = p00 * cm + p11 * cd + p22 * cm;
    b[5] = gj[1] * gk[2] * co;
    b[6] = gj[2] * gk[0] * co;
    b[7] = gj[2] * gk[1] * co;
    b[8] = p00 * cm + p11 * cm + p22 * cd;
}

template <typename T>
__device__ __forceinline__ T lap_entry(const T *gj, const T *gk)
{
    return gj[0] * gk[0] + gj[1] * gk[1] + gj[2] * gk[2];
}

// --------------------------------------------------------------------------------
// 3x3 inverse in the order of reference LAPACK DGETRF (DGETF2: IDAMAX pivot, row swap,
// reciprocal scaling, DGER update) followed by DGETRI (DTRTI2 inverse of U, DGEMV
// column sweep, column interchanges). a is column-major, A(r,c) = a[c*3+r]. All
// indices are compile-time after unrolling; pivot swaps use selects so the matrix
// stays in registers. Returns LAPACK's info (0 = ok); a singular U leaves the LU
// factors in a, as DGETRI does.
// --------------------------------------------------------------------------------
__device__ __forceinline__ void swap_rows_sel(double (&a)[9], int j, int jp)
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i <= j) continue;
        const bool s = (jp == i);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double x = a[c * 3 + j], y = a[c * 3 + i];
            a[c * 3 + j] = s ? y : x;
            a[c * 3 + i] = s ? x : y;
        }
    }
}

__device__ __forceinline__ void swap_cols_sel(double (&a)[9], int j, int jp)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (c <= j) continue;
        const bool s = (jp == c);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double x = a[j * 3 + i], y = a[c * 3 + i];
            a[j * 3 + i] = s ? y : x;
            a[c * 3 + i] = s ? x : y;
        }
    }
}

__device__ __forceinline__ int inv3_lapack(double (&a)[9])
{
    int info = 0;
    int ipiv[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        int jp = j;
        double amax = fabs(a[j * 3 + j]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (i <= j) continue;
            const double v = fabs(a[j * 3 + i]);
            if (v > amax) { amax = v; jp = i; }
        }
        ipiv[j] = jp;
        // pivot value after the (possible) swap
        double piv = a[j * 3 + j];
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (i > j && jp == i) piv = a[j * 3 + i];
        if (piv != 0.0) {
            swap_rows_sel(a, j, jp);
            if (j < 2) {
                if (fabs(a[j * 3 + j]) >= 2.2250738585072014e-308) {
                    const double r = 1.0 / a[j * 3 + j];
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        if (i > j) a[j * 3 + i] *= r;
                } else {
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        if (i > j) a[j * 3 + i] = a[j * 3 + i] / a[j * 3 + j];
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < 2) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (c <= j) continue;
                if (a[c * 3 + j] != 0.0) {
                    const double t = -a[c * 3 + j];
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        if (i > j) a[c * 3 + i] = a[c * 3 + i] + a[j * 3 + i] * t;
                }
            }
        }
    }
    if (a[0] == 0.0 || a[4] == 0.0 || a[8] == 0.0) return info ? info : 1;
    // DTRTI2, upper, non-unit
#pragma unroll
    for (int j = 0; j < 3; j++) {
        a[j * 3 + j] = 1.0 / a[j * 3 + j];
        const double ajj = -a[j * 3 + j];
#pragma unroll
        for (int j2 = 0; j2 < 3; j2++) {
            if (j2 >= j) continue;
            if (a[j * 3 + j2] != 0.0) {
                const double t = a[j * 3 + j2];
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (i < j2) a[j * 3 + i] = a[j * 3 + i] + t * a[j2 * 3 + i];
                a[j * 3 + j2] = a[j * 3 + j2] * a[j2 * 3 + j2];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (i < j) a[j * 3 + i] *= ajj;
    }
    // DGETRI column sweep
    double work[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 2; j >= 0; j--) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (i > j) { work[i] = a[j * 3 + i]; a[j * 3 + i] = 0.0; }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (c <= j) continue;
            const double t = -work[c];
#pragma unroll
            for (int i = 0; i < 3; i++) a[j * 3 + i] = a[j * 3 + i] + t * a[c * 3 + i];
        }
    }
#pragma unroll
    for (int j = 1; j >= 0; j--) swap_cols_sel(a, j, ipiv[j]);
    return info;
}

// Block-Jacobi inverse of one node's ela diagonal block, src/Fortran/elasclpr.f:19-53:
// blk is the node's 9 values in memory order; Fortran reads them as prec(ki,kj) =
// blk[kj*3+ki]. Dofs with mask bit k set get row k and column k zeroed and a unit
// diagonal; then, when the row holds its diagonal entry, DGETRF + DGETRI.
__device__ __forceinline__ int ela_invert_block(double (&blk)[9], int mask, bool hasDiag)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (mask & (1 << k)) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                blk[q * 3 + k] = 0.0;
                blk[k * 3 + q] = 0.0;
            }
            blk[k * 3 + k] = 1.0;
        }
    }
    if (!hasDiag) return 0;
    return inv3_lapack(blk);
}

}  // namespace mf
