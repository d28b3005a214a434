// Device-side mesh setup, bit-exact twins of the host setup (src/matrix.cc:25-91 and
// DC_create_nodeToElem as called at src/main.cc:243-250): for meshes that arrive at a
// server already resident on the GPU. Greedy first-fit coloring
// (src/coloring.cc:84-109) is a sequential scan over elements and stays on the host.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>

#include "device_common.h"
#include "minifem.h"

namespace {

__global__ void k_count_nodes(const int *__restrict__ e2n, int64_t n4, int N, int *__restrict__ cnt,
                              int *__restrict__ bad)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n4) return;
    const int v = e2n[q] - 1;
    if (v < 0 || v >= N) {
        atomicOr(bad, 1);
        return;
    }
    atomicAdd(cnt + v, 1);
}

// index[] holds the exclusive scan; cur[] starts as a copy of it
__global__ void k_fill_nodes(const int *__restrict__ e2n, int64_t n4, int *__restrict__ cur,
                             int *__restrict__ value)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n4) return;
    const int v = e2n[q] - 1;
    value[atomicAdd(cur + v, 1)] = (int)(q >> 2);
}

// each node's element list in ascending order (insertion sort: lists are short)
__global__ void k_sort_lists(const int *__restrict__ index, int N, int *__restrict__ value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int b = index[i], e = index[i + 1];
    for (int a = b + 1; a < e; a++) {
        const int x = value[a];
        int c = a - 1;
        while (c >= b && value[c] > x) {
            value[c + 1] = value[c];
            c--;
        }
        value[c + 1] = x;
    }
}

// Row i in first-seen order: node i's elements ascending, each element's vertices in
// order, a column appended when not seen yet (src/matrix.cc:67-85). With out == nullptr
// only the length is counted.
__device__ int row_columns(int i, const int *__restrict__ index, const int *__restrict__ value,
                           const int *__restrict__ e2n, int *out, int cap)
{
    int len = 0;
    for (int j = index[i]; j < index[i + 1]; j++) {
        const int el = value[j];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int c = e2n[(int64_t)el * 4 + k];  // 1-based, as col[] stores it
            bool seen = false;
            if (out) {
                for (int q = 0; q < len; q++) seen |= out[q] == c;
            } else {
                // count pass: compare against the earlier vertices of earlier elements
                for (int j2 = index[i]; j2 <= j && !seen; j2++) {
                    const int el2 = value[j2];
                    const int kmax = j2 == j ? k : 4;
                    for (int k2 = 0; k2 < kmax; k2++) seen |= e2n[(int64_t)el2 * 4 + k2] == c;
                }
            }
            if (!seen) {
                if (out && len < cap) out[len] = c;
                len++;
            }
        }
    }
    return len;
}

__global__ void k_row_len(const int *__restrict__ index, const int *__restrict__ value,
                          const int *__restrict__ e2n, int N, int *__restrict__ len)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    len[i] = row_columns(i, index, value, e2n, nullptr, 0);
}

__global__ void k_row_fill(const int *__restrict__ index, const int *__restrict__ value,
                           const int *__restrict__ e2n, const int *__restrict__ row, int N,
                           int *__restrict__ col)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    row_columns(i, index, value, e2n, col + row[i], row[i + 1] - row[i]);
}

// elemToEdge (src/matrix.cc:25-52): for each element, the 16 CSR positions of its
// (i, j) vertex pairs, row-major
__global__ void k_elem_to_edge(const int *__restrict__ row, const int *__restrict__ col,
                               const int *__restrict__ e2n, int nE, int *__restrict__ e2e)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nE * 16) return;
    const int64_t el = t >> 4;
    const int i = (int)((t >> 2) & 3), j = (int)(t & 3);
    const int a = e2n[el * 4 + i] - 1, c = e2n[el * 4 + j];
    int pos = -1;
    for (int q = row[a]; q < row[a + 1]; q++)
        if (col[q] == c) {
            pos = q;
            break;
        }
    e2e[t] = pos;
}

int scan_exclusive(const int *in, int *out, int n, hipStream_t s)
{
    size_t tmp = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s) != hipSuccess) return MF_ERR_HIP;
    void *buf = nullptr;
    if (hipMalloc(&buf, std::max<size_t>(tmp, 16)) != hipSuccess) return MF_ERR_NOMEM;
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(buf, tmp, in, out, n, s);
    (void)hipFree(buf);
    return e == hipSuccess ? MF_OK : MF_ERR_HIP;
}

inline dim3 grid_for(int64_t n, int bt) { return dim3((unsigned)((n + bt - 1) / bt)); }

}  // namespace

extern "C" int mf_dev_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes, int *index,
                                        int *value, void *stream)
{
    if (nbElem < 0 || nbNodes < 0 || !index || (nbElem > 0 && (!elemToNode || !value)))
        return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t n4 = (int64_t)nbElem * 4;
    int *cnt = nullptr, *bad = nullptr;
    if (hipMalloc(&cnt, ((size_t)nbNodes + 2) * sizeof(int)) != hipSuccess) return MF_ERR_NOMEM;
    bad = cnt + nbNodes + 1;
    int st = MF_OK;
    if (hipMemsetAsync(cnt, 0, ((size_t)nbNodes + 2) * sizeof(int), s) != hipSuccess) st = MF_ERR_HIP;
    if (st == MF_OK && n4 > 0) {
        hipLaunchKernelGGL(k_count_nodes, grid_for(n4, 256), dim3(256), 0, s, elemToNode, n4,
                           nbNodes, cnt, bad);
        if (hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
    }
    int hbad = 0;
    if (st == MF_OK && hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_OK && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    if (st == MF_OK && hbad) st = MF_ERR_ARG;  // a node id outside 1..N
    if (st == MF_OK) st = scan_exclusive(cnt, index, nbNodes + 1, s);
    if (st == MF_OK && n4 > 0) {
        // cnt becomes the fill cursor
        if (hipMemcpyAsync(cnt, index, (size_t)nbNodes * sizeof(int), hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
            st = MF_ERR_HIP;
        if (st == MF_OK) {
            hipLaunchKernelGGL(k_fill_nodes, grid_for(n4, 256), dim3(256), 0, s, elemToNode, n4,
                               cnt, value);
            hipLaunchKernelGGL(k_sort_lists, grid_for(nbNodes, 256), dim3(256), 0, s, index,
                               nbNodes, value);
            if (hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == MF_OK) st = MF_ERR_HIP;
    (void)hipFree(cnt);
    return st;
}

extern "C" int mf_dev_create_nodeToNode(int *row, int *col, const int *n2eIndex,
                                        const int *n2eValue, const int *elemToNode, int nbNodes,
                                        int nbEdges, void *stream)
{
    if (!row || nbNodes < 0 || nbEdges < 0 || (nbNodes > 0 && (!n2eIndex || !elemToNode)))
        return MF_ERR_ARG;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int *len = nullptr;
    if (hipMalloc(&len, ((size_t)nbNodes + 1) * sizeof(int)) != hipSuccess) return MF_ERR_NOMEM;
    int st = MF_OK;
    if (hipMemsetAsync(len + nbNodes, 0, sizeof(int), s) != hipSuccess) st = MF_ERR_HIP;
    if (st == MF_OK && nbNodes > 0) {
        hipLaunchKernelGGL(k_row_len, grid_for(nbNodes, 128), dim3(128), 0, s, n2eIndex, n2eValue,
                           elemToNode, nbNodes, len);
        if (hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
    }
    if (st == MF_OK) st = scan_exclusive(len, row, nbNodes + 1, s);
    int nnz = 0;
    if (st == MF_OK &&
        hipMemcpyAsync(&nnz, row + nbNodes, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_OK && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    if (st == MF_OK && nnz > nbEdges) st = MF_ERR_NNZ;  // the pattern does not fit col[]
    if (st == MF_OK && nbNodes > 0 && nnz > 0) {
        if (!col) st = MF_ERR_ARG;
        else {
            hipLaunchKernelGGL(k_row_fill, grid_for(nbNodes, 128), dim3(128), 0, s, n2eIndex,
                               n2eValue, elemToNode, row, nbNodes, col);
            if (hipGetLastError() != hipSuccess) st = MF_ERR_HIP;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == MF_OK) st = MF_ERR_HIP;
    (void)hipFree(len);
    return st == MF_OK ? nnz : st;
}

extern "C" int mf_dev_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                                        int *elemToEdge, int nbElem, void *stream)
{
    if (nbElem < 0 || (nbElem > 0 && (!row || !col || !elemToNode || !elemToEdge)))
        return MF_ERR_ARG;
    if (nbElem == 0) return MF_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_elem_to_edge, grid_for((int64_t)nbElem * 16, 256), dim3(256), 0, s, row,
                       col, elemToNode, nbElem, elemToEdge);
    if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? MF_OK : MF_ERR_HIP;
}
