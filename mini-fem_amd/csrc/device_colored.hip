// Drop-in GPU twins of the reference's per-iteration entry points:
//   assembly()        -> mf_assembly / mf_assembly_ex   (COLORING semantics)
//   prec_init()       -> mf_prec_init
//   prec_inversion()  -> mf_prec_inversion
//   compute_double_norm() -> mf_norm
// The assembly is one launch per color with 16 lanes per element (one per element block):
// elements of one color share no node, so their scatter-adds never touch the same CSR
// entry and each pass is race-free and deterministic; the per-entry summation order is
// the color order, i.e. the reference COLORING build's element order.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "device_common.h"
#include "minifem.h"

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_fill(double *__restrict__ p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// CSR position of the (j,k) pair of one element: elemToEdge when present, else the
// reference's row search (src/assembly.cc:415-421).
__device__ __forceinline__ int elem_position(const int *__restrict__ row,
                                             const int *__restrict__ col,
                                             const int *__restrict__ e2e, int e, int nj, int nk,
                                             int q)
{
    if (e2e) return e2e[(int64_t)e * 16 + q];
    for (int l = row[nj - 1]; l < row[nj]; l++)
        if (col[l] == nk) return l;
    return -1;
}

// One color: 16 lanes per element, lane q = 4j + k adds the element's (j,k) block into
// its CSR entry. Elements of one color share no node, so no two lanes of a launch touch
// the same entry; each entry receives its elements' blocks in color order, as in the
// reference COLORING build. The 16 lanes of an element compute its gradients alike.
template <int OP, int PREC>
__global__ void __launch_bounds__(kBlock)
    k_assemble_color(const double *__restrict__ coord, const double *__restrict__ coef,
                     double *values, const int *__restrict__ row,
                     const int *__restrict__ col, const int4 *__restrict__ e2n,
                     const int *__restrict__ e2e, int first, int count)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)count * 16) return;
    const int e = first + (int)(t >> 4), q = (int)(t & 15), j = q >> 2, k = q & 3;
    const int4 v = e2n[e];
    double c[4][3], g[4][3];
    mf::load_coords(coord, v, c);
    mf::elem_grad(c, g);
    const int n[4] = {v.x, v.y, v.z, v.w};
    const int p = elem_position(row, col, e2e, e, n[j], n[k], q);
    if (p < 0) return;
    if (OP == MF_OP_LAP) {
        const double kd = coef ? coef[e] : 1.0;
        if (PREC == MF_PREC_FP64) {
            double x = mf::lap_entry(g[j], g[k]);
            if (coef) x = kd * x;
            values[p] += x;
        } else {
            float gj[3], gk[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gj[d] = (float)g[j][d];
                gk[d] = (float)g[k][d];
            }
            const float x = (float)kd * mf::lap_entry(gj, gk);
            values[p] += (double)x;
        }
    } else {
        const mf::ElaCoef kc = mf::ela_coef(coef, e);
        double *out = values + (int64_t)p * 9;
        if (PREC == MF_PREC_FP64) {
            double b[9];
            mf::ela_block(g[j], g[k], kc.cd, kc.co, kc.cm, b);
#pragma unroll
            for (int r = 0; r < 9; r++) out[r] += b[r];
        } else {
            float gj[3], gk[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                gj[d] = (float)g[j][d];
                gk[d] = (float)g[k][d];
            }
            float b[9];
            mf::ela_block(gj, gk, (float)kc.cd, (float)kc.co, (float)kc.cm, b);
#pragma unroll
            for (int r = 0; r < 9; r++) out[r] += (double)b[r];
        }
    }
}

// prec_init, src/preconditioner.cc:52-87: zero, then copy the diagonal entry's d values.
__global__ void __launch_bounds__(kBlock)
    k_prec_init(double *__restrict__ prec, const double *__restrict__ values,
                const int *__restrict__ row, const int *__restrict__ col, int N, int dim)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double *out = prec + (int64_t)i * dim;
    for (int k = 0; k < dim; k++) out[k] = 0.0;
    for (int l = row[i]; l < row[i + 1]; l++) {
        if (col[l] - 1 == i) {
            const double *src = values + (int64_t)l * dim;
            for (int k = 0; k < dim; k++) out[k] = src[k];
            break;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_prec_inv_lap(double *__restrict__ prec, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) prec[i] = 1.0 / prec[i];
}

// prec_inversion for ela: one ela_invert_prec per node (src/preconditioner.cc:44-46).
__global__ void __launch_bounds__(kBlock)
    k_prec_inv_ela(double *__restrict__ prec, const int *__restrict__ row,
                   const int *__restrict__ col, const int *__restrict__ cb, int N,
                   int *__restrict__ singular)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double blk[9];
    double *p = prec + (int64_t)i * 9;
#pragma unroll
    for (int q = 0; q < 9; q++) blk[q] = p[q];
    int mask = 0;
    if (cb) {
        mask = (cb[i] != 0 ? 1 : 0) | (cb[(int64_t)N + i] != 0 ? 2 : 0) |
               (cb[2 * (int64_t)N + i] != 0 ? 4 : 0);
    }
    bool hasDiag = false;
    for (int l = row[i]; l < row[i + 1]; l++)
        if (col[l] == i + 1) { hasDiag = true; break; }
    const int info = mf::ela_invert_block(blk, mask, hasDiag);
#pragma unroll
    for (int q = 0; q < 9; q++) p[q] = blk[q];
    if (info != 0 && singular) atomicAdd(singular, 1);
}

// Deterministic sum of squares: fixed grid, fixed per-thread stride order, fixed LDS
// tree; then one block folds the partials in index order.
constexpr int kNormBlocks = 1024;

__global__ void __launch_bounds__(kBlock)
    k_sumsq_partial(const double *__restrict__ x, int64_t n, double *__restrict__ part)
{
    __shared__ double s[kBlock];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)kNormBlocks * kBlock)
        acc += x[i] * x[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kBlock) k_sumsq_final(double *__restrict__ part)
{
    __shared__ double s[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < kNormBlocks; i += kBlock) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[0] = s[0];
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <int OP, int PREC>
int launch_colors(const double *coord, const double *coef, double *values, const int *row,
                  const int *col, const int *e2n, const int *e2e, const int *colorToElem,
                  int nbColors, hipStream_t s)
{
    for (int c = 0; c < nbColors; c++) {
        const int first = colorToElem[c], count = colorToElem[c + 1] - first;
        if (count <= 0) continue;
        hipLaunchKernelGGL((k_assemble_color<OP, PREC>), dim3(grid_for((int64_t)count * 16)),
                           dim3(kBlock),
                           0, s, coord, coef, values, row, col,
                           reinterpret_cast<const int4 *>(e2n), e2e, first, count);
        MF_CHECK_LAUNCH();
    }
    return MF_OK;
}

}  // namespace

// Scratch of the allocation-free norm (mf_norm_noalloc): the kNormBlocks per-workgroup
// partial sums of k_sumsq_partial, folded in place by k_sumsq_final. 8 KiB, held by
// every GPU the library runs on until the process ends, each GPU its own copy at its own
// address. Host code never passes `scratch` itself (a host-side stand-in): it uses the
// per-device address resolved by hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[kNormBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device `dev`, resolved once per device
// (at mf_device_init, or on the first allocation-free call on a device that skipped it).
int norm_scratch(int dev, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return MF_ERR_HIP;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MF_ERR_HIP;
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return MF_OK;
}

}  // namespace

extern "C" const char *mf_status_string(int status)
{
    switch (status) {
    case MF_OK: return "ok";
    case MF_ERR_ARG: return "invalid argument";
    case MF_ERR_IO: return "file missing, truncated or unwritable";
    case MF_ERR_COLORS: return "Not enough colors (more than 128)";
    case MF_ERR_NNZ: return "CSR pattern larger than the nnz capacity";
    case MF_ERR_HIP: return "HIP runtime error or no GPU";
    case MF_ERR_NOMEM: return "allocation failed";
    case MF_ERR_BUDGET: return "request exceeds the HBM budget";
    case MF_ERR_STATE: return "object used in the wrong state";
    default: return "unknown status";
    }
}

extern "C" int mf_version(void) { return 1; }

extern "C" int mf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return MF_ERR_HIP;
    return n;
}

extern "C" int mf_device_init(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MF_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return MF_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MF_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::fprintf(stderr, "minifem: device %d is %s, this library is built for gfx950\n",
                     device, prop.gcnArchName);
        return MF_ERR_HIP;
    }
    double *unused = nullptr;
    return norm_scratch(device, &unused);
}

extern "C" int mf_assembly_ex(const double *coord, const double *coef, double *values,
                              const int *row, const int *col, const int *elemToNode,
                              const int *elemToEdge, int nbElem, int nbEdges,
                              int operatorDim, int operatorID, const int *colorToElem,
                              int nbColors, int precision, void *stream)
{
    if (nbElem < 0 || nbEdges < 0 || nbColors < 0 || nbColors > MF_MAX_COLOR) return MF_ERR_ARG;
    if (!(operatorID == MF_OP_LAP && operatorDim == 1) &&
        !(operatorID == MF_OP_ELA && operatorDim == 9))
        return MF_ERR_ARG;
    if (precision != MF_PREC_FP64 && precision != MF_PREC_MIXED) return MF_ERR_ARG;
    if (!values || (nbElem > 0 && (!coord || !elemToNode || !colorToElem || !row ||
                                   (!elemToEdge && !col))))
        return MF_ERR_ARG;
    if (nbElem > 0 && (colorToElem[0] != 0 || colorToElem[nbColors] != nbElem)) return MF_ERR_ARG;
    for (int c = 0; c < nbColors; c++)
        if (colorToElem[c + 1] < colorToElem[c]) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t nv = (int64_t)nbEdges * operatorDim;
    if (nv > 0) {
        const int64_t g = grid_for(nv);
        hipLaunchKernelGGL(k_fill, dim3((unsigned)(g < 8192 ? g : 8192)), dim3(kBlock), 0, s,
                           values, nv, 0.0);
        MF_CHECK_LAUNCH();
    }
    if (nbElem == 0) return MF_OK;
    if (operatorID == MF_OP_LAP)
        return precision == MF_PREC_FP64
                   ? launch_colors<MF_OP_LAP, MF_PREC_FP64>(coord, coef, values, row, col,
                                                            elemToNode, elemToEdge,
                                                            colorToElem, nbColors, s)
                   : launch_colors<MF_OP_LAP, MF_PREC_MIXED>(coord, coef, values, row, col,
                                                             elemToNode, elemToEdge,
                                                             colorToElem, nbColors, s);
    return precision == MF_PREC_FP64
               ? launch_colors<MF_OP_ELA, MF_PREC_FP64>(coord, coef, values, row, col,
                                                        elemToNode, elemToEdge, colorToElem,
                                                        nbColors, s)
               : launch_colors<MF_OP_ELA, MF_PREC_MIXED>(coord, coef, values, row, col,
                                                         elemToNode, elemToEdge, colorToElem,
                                                         nbColors, s);
}

extern "C" int mf_assembly(const double *coord, double *values, const int *row,
                           const int *col, const int *elemToNode, const int *elemToEdge,
                           int nbElem, int nbEdges, int operatorDim, int operatorID,
                           const int *colorToElem, int nbColors, void *stream)
{
    return mf_assembly_ex(coord, nullptr, values, row, col, elemToNode, elemToEdge, nbElem,
                          nbEdges, operatorDim, operatorID, colorToElem, nbColors,
                          MF_PREC_FP64, stream);
}

extern "C" int mf_prec_init(double *prec, const double *values, const int *row,
                            const int *col, int nbNodes, int operatorDim, void *stream)
{
    if (nbNodes < 0 || (operatorDim != 1 && operatorDim != 9)) return MF_ERR_ARG;
    if (nbNodes == 0) return MF_OK;
    if (!prec || !values || !row || !col) return MF_ERR_ARG;
    hipLaunchKernelGGL(k_prec_init, dim3(grid_for(nbNodes)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), prec, values, row, col, nbNodes,
                       operatorDim);
    MF_CHECK_LAUNCH();
    return MF_OK;
}

extern "C" int mf_prec_inversion(double *prec, const int *row, const int *col,
                                 const int *checkBounds, int nbNodes, int operatorID,
                                 int *singular, void *stream)
{
    if (nbNodes < 0 || (operatorID != MF_OP_LAP && operatorID != MF_OP_ELA)) return MF_ERR_ARG;
    if (nbNodes == 0) return MF_OK;
    if (!prec) return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (operatorID == MF_OP_LAP) {
        hipLaunchKernelGGL(k_prec_inv_lap, dim3(grid_for(nbNodes)), dim3(kBlock), 0, s, prec,
                           nbNodes);
    } else {
        if (!row || !col) return MF_ERR_ARG;
        hipLaunchKernelGGL(k_prec_inv_ela, dim3(grid_for(nbNodes)), dim3(kBlock), 0, s, prec,
                           row, col, checkBounds, nbNodes, singular);
    }
    MF_CHECK_LAUNCH();
    return MF_OK;
}

extern "C" int mf_norm(const double *x, int64_t n, double *result, void *stream)
{
    if (n < 0 || !result || (n > 0 && !x)) return MF_ERR_ARG;
    if (n == 0) { *result = 0.0; return MF_OK; }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double *part = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&part), kNormBlocks * sizeof(double), s) !=
        hipSuccess)
        return MF_ERR_NOMEM;
    hipLaunchKernelGGL(k_sumsq_partial, dim3(kNormBlocks), dim3(kBlock), 0, s, x, n, part);
    hipLaunchKernelGGL(k_sumsq_final, dim3(1), dim3(kBlock), 0, s, part);
    double h = 0.0;
    hipError_t e1 = hipMemcpyAsync(&h, part, sizeof(double), hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipStreamSynchronize(s);
    hipError_t e3 = hipFreeAsync(part, s);
    hipError_t e4 = hipStreamSynchronize(s);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess)
        return MF_ERR_HIP;
    *result = __builtin_sqrt(h);
    return MF_OK;
}

extern "C" int mf_norm_noalloc(const double *x, int64_t n, double *result, void *stream)
{
    if (n < 0 || !result || (n > 0 && !x)) return MF_ERR_ARG;
    if (n == 0) { *result = 0.0; return MF_OK; }
    int dev = 0;
    double *part = nullptr;
    if (hipGetDevice(&dev) != hipSuccess) return MF_ERR_HIP;
    const int st = norm_scratch(dev, &part);
    if (st != MF_OK) return st;
    // the same two kernels and order as mf_norm: every workgroup writes its own partial
    // (zero when it sums nothing), so nothing in the scratch has to be cleared first
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_sumsq_partial, dim3(kNormBlocks), dim3(kBlock), 0, s, x, n, part);
    hipLaunchKernelGGL(k_sumsq_final, dim3(1), dim3(kBlock), 0, s, part);
    double h = 0.0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&h, part, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MF_ERR_HIP;
    *result = __builtin_sqrt(h);
    return MF_OK;
}
