// Device-side mesh setup, bit-exact twins of the host setup (src/matrix.cc:25-91,
// DC_create_nodeToElem as called at src/main.cc:243-250, and the greedy coloring of
// src/coloring.cc:46-109): for meshes that arrive at a server already resident on the GPU.
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

// ---- greedy first-fit coloring (src/coloring.cc:46-81), exact, level by level ----
//
// Element e takes the lowest color that no lower-index element sharing a node with it
// holds. At node v the elements are listed in ascending order (nodeToElem), so e's color
// depends on v's list only up to e. Once every element before e in each of its nodes'
// lists is colored, e's color is fixed: the first color missing from the OR of its nodes'
// color masks, each mask holding the colors of the elements colored at that node so far
// (the host scan's per-node masks). Elements become ready in levels: level 0 has no lower
// neighbour at all, and an element is ready once its predecessor at each node is colored.
// Two elements ready in the same level never share a node (the later one's predecessor
// at a shared node would be uncolored), so one level colors its elements in parallel, and
// the result is the sequential scan's, bit for bit. The 100^3 cube needs 2,750 levels of
// about 2,100 elements each, one launch per level.

// Per element: at each distinct node, the element that follows it in the node's list
// (-1: none), and the number of distinct nodes where an element precedes it.
__global__ void k_color_links(const int4 *__restrict__ e2n, int nE, const int *__restrict__ index,
                              const int *__restrict__ value, int4 *__restrict__ next,
                              int *__restrict__ npre)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int4 q = e2n[e];
    const int v[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
    int nx[4] = {-1, -1, -1, -1}, np = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        bool dup = false;
#pragma unroll
        for (int j = 0; j < k; j++) dup |= v[j] == v[k];
        if (dup) continue;  // a node the element repeats counts once
        const int b = index[v[k]], en = index[v[k] + 1];
        int lo = b, hi = en;  // first entry >= e
        while (lo < hi) {
            const int m = (lo + hi) >> 1;
            if (value[m] < e) lo = m + 1;
            else hi = m;
        }
        int last = lo;  // past e's entries (two when the element repeats the node)
        while (last < en && value[last] == e) last++;
        np += lo > b;
        if (last < en) nx[k] = value[last];
    }
    next[e] = make_int4(nx[0], nx[1], nx[2], nx[3]);
    npre[e] = np;
}

__global__ void k_color_seed(const int *__restrict__ npre, int nE, int *__restrict__ front,
                             int *__restrict__ count)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nE && npre[e] == 0) front[atomicAdd(count, 1)] = e;
}

// One level per launch: launches on one stream run in order, which orders the levels, and
// each level spreads over kLevelBlocks workgroups (a level's cost is its chain of memory
// round trips, so more CUs keep more of them in flight). Level L reads its list from
// front[L & 1] and its length from counts[L % 3], appends the next level to
// front[(L + 1) & 1] and counts[(L + 1) % 3], and clears counts[(L + 2) % 3] for level L + 1.
// status: [1] elements colored, [2] nonzero when an element needs more than MF_MAX_COLOR
// colors.
constexpr int kLevelThreads = 256, kLevelBlocks = 32, kLevelsPerCheck = 64;

__global__ void __launch_bounds__(kLevelThreads)
k_color_level(const int4 *__restrict__ e2n, const int4 *__restrict__ next,
              const int *__restrict__ npre, int *__restrict__ cnt, const int *__restrict__ cur,
              int *__restrict__ nxt, int *__restrict__ counts, int L, ulonglong2 *masks,
              int *__restrict__ part, int *__restrict__ status)
{
    const int n = counts[L % 3];
    const int gt = blockIdx.x * kLevelThreads + threadIdx.x;
    if (gt == 0) {
        counts[(L + 2) % 3] = 0;
        if (n > 0) atomicAdd(status + 1, n);
    }
    for (int i = gt; i < n; i += kLevelBlocks * kLevelThreads) {
        const int e = cur[i];
        const int4 q = e2n[e], nq = next[e];
        const int v[4] = {q.x - 1, q.y - 1, q.z - 1, q.w - 1};
        const int c[4] = {nq.x, nq.y, nq.z, nq.w};
        ulonglong2 m[4];
        int old[4], np[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            m[k] = masks[v[k]];
            old[k] = c[k] >= 0 ? atomicAdd(cnt + c[k], 1) : 0;
            np[k] = c[k] >= 0 ? npre[c[k]] : 0;
        }
        const uint64_t lo = m[0].x | m[1].x | m[2].x | m[3].x;
        const uint64_t hi = m[0].y | m[1].y | m[2].y | m[3].y;
        int color = ~lo ? __builtin_ctzll(~lo) : ~hi ? 64 + __builtin_ctzll(~hi) : 128;
        if (color >= MF_MAX_COLOR) {
            atomicOr(status + 2, 1);
            color = MF_MAX_COLOR - 1;  // keeps the levels going; the call fails
        }
        part[e] = color;
        const uint64_t blo = color < 64 ? 1ull << color : 0ull;
        const uint64_t bhi = color < 64 ? 0ull : 1ull << (color - 64);
        // no other element of this level holds these nodes (a repeated node gets the same
        // value twice)
#pragma unroll
        for (int k = 0; k < 4; k++) masks[v[k]] = make_ulonglong2(m[k].x | blo, m[k].y | bhi);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            // no successor at this node, or c[k] still waits for another node
            if (c[k] < 0 || old[k] + 1 != np[k]) continue;
            nxt[atomicAdd(counts + (L + 1) % 3, 1)] = c[k];
        }
    }
}

__global__ void k_iota(int *__restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Elements sorted by (color, index): position i holds element order[i], whose new index
// is i (the stable counting sort of src/coloring.cc:27-42 and DC_create_permutation).
// Each color's first position goes to colorToElem.
__global__ void k_color_perm(const int *__restrict__ key, const int *__restrict__ order, int nE,
                             int *__restrict__ perm, int *__restrict__ c2e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nE) return;
    perm[order[i]] = i;
    if (i == 0 || key[i] != key[i - 1]) c2e[key[i]] = i;
    if (i == nE - 1) c2e[key[i] + 1] = nE;
}

__global__ void k_check_perm(const int *__restrict__ perm, int n, int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (perm[i] < 0 || perm[i] >= n)) atomicOr(bad, 1);
}

// row i of src to row perm[i] of dst
template <typename T>
__global__ void k_permute_rows(const T *__restrict__ src, const int *__restrict__ perm, int n,
                               int dim, T *__restrict__ dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * dim) return;
    const int64_t i = t / dim, d = t - i * dim;
    dst[(int64_t)perm[i] * dim + d] = src[t];
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
    template <typename T>
    T *get() const { return static_cast<T *>(p); }
};

// Waits for the stream when it goes out of scope: declared after the DevBufs of a call, it
// runs before they free their memory, on every return path.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

inline int dev_alloc(DevBuf &b, size_t bytes)
{
    return hipMalloc(&b.p, std::max<size_t>(bytes, 16)) == hipSuccess ? MF_OK : MF_ERR_NOMEM;
}

}  // namespace

extern "C" int mf_dev_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                                        int nbElem, int nbNodes, int *colorToElem,
                                        int *nbColors, void *stream)
{
    if (nbElem < 0 || nbNodes < 0 || !colorPerm || !colorToElem || !nbColors ||
        (nbElem > 0 && !elemToNode) || (reinterpret_cast<uintptr_t>(elemToNode) & 15))
        return MF_ERR_ARG;  // elements are read as 16-byte rows
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (nbElem == 0) {
        *nbColors = 0;
        const int z = 0;
        if (hipMemcpyAsync(colorToElem, &z, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return MF_ERR_HIP;
        return MF_OK;
    }
    const size_t nE = (size_t)nbElem, N = (size_t)nbNodes;
    DevBuf index, value, next, npre, cnt, fa, fb, masks, part, misc, keys, order, tmp;
    const StreamDrain drain{s};  // nothing in flight when the buffers are freed
    int st = MF_OK;
    if (st == MF_OK) st = dev_alloc(index, (N + 1) * sizeof(int));
    if (st == MF_OK) st = dev_alloc(value, 4 * nE * sizeof(int));
    if (st == MF_OK) st = dev_alloc(next, nE * sizeof(int4));
    if (st == MF_OK) st = dev_alloc(npre, nE * sizeof(int));
    if (st == MF_OK) st = dev_alloc(cnt, nE * sizeof(int));
    if (st == MF_OK) st = dev_alloc(fa, nE * sizeof(int));
    if (st == MF_OK) st = dev_alloc(fb, nE * sizeof(int));
    if (st == MF_OK) st = dev_alloc(masks, N * sizeof(ulonglong2));
    if (st == MF_OK) st = dev_alloc(misc, 6 * sizeof(int));  // counts[3], status[3]
    if (st == MF_OK && !colorPart) st = dev_alloc(part, nE * sizeof(int));
    if (st != MF_OK) return st;
    int *partp = colorPart ? colorPart : part.get<int>();
    // nodeToElem (each node's elements ascending), as the host's
    st = mf_dev_create_nodeToElem(elemToNode, nbElem, nbNodes, index.get<int>(), value.get<int>(),
                                  stream);
    if (st != MF_OK) return st;
    const int4 *e2n = reinterpret_cast<const int4 *>(elemToNode);
    int *counts = misc.get<int>(), *status = counts + 3;
    if (hipMemsetAsync(cnt.p, 0, nE * sizeof(int), s) != hipSuccess ||
        hipMemsetAsync(masks.p, 0, N * sizeof(ulonglong2), s) != hipSuccess ||
        hipMemsetAsync(misc.p, 0, 6 * sizeof(int), s) != hipSuccess)
        return MF_ERR_HIP;
    hipLaunchKernelGGL(k_color_links, grid_for(nbElem, 256), dim3(256), 0, s, e2n, nbElem,
                       index.get<int>(), value.get<int>(), next.get<int4>(), npre.get<int>());
    hipLaunchKernelGGL(k_color_seed, grid_for(nbElem, 256), dim3(256), 0, s, npre.get<int>(),
                       nbElem, fa.get<int>(), counts);
    if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
    // levels in batches of kLevelsPerCheck launches, then a look at the next level's
    // length: launches past the last level find an empty list and return at once
    int *front[2] = {fa.get<int>(), fb.get<int>()};
    int L = 0, remaining = 1;
    while (remaining > 0) {
        if (L > nbElem) return MF_ERR_ARG;  // more levels than elements cannot happen
        for (int b = 0; b < kLevelsPerCheck; b++, L++)
            hipLaunchKernelGGL(k_color_level, dim3(kLevelBlocks), dim3(kLevelThreads), 0, s, e2n,
                               next.get<int4>(), npre.get<int>(), cnt.get<int>(), front[L & 1],
                               front[(L + 1) & 1], counts, L, masks.get<ulonglong2>(), partp,
                               status);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&remaining, counts + L % 3, sizeof(int), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return MF_ERR_HIP;
    }
    int hs[3] = {0, 0, 0};
    if (hipMemcpyAsync(hs, status, sizeof(hs), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MF_ERR_HIP;
    if (hs[2]) return MF_ERR_COLORS;
    if (hs[1] != nbElem) return MF_ERR_ARG;  // not every element became ready
    // stable sort of the element ids by color: 7 key bits (colors < 128)
    if (dev_alloc(keys, nE * sizeof(int)) != MF_OK || dev_alloc(order, 2 * nE * sizeof(int)) != MF_OK)
        return MF_ERR_NOMEM;
    int *kout = keys.get<int>(), *vin = order.get<int>(), *vout = vin + nE;
    hipLaunchKernelGGL(k_iota, grid_for(nbElem, 256), dim3(256), 0, s, vin, nbElem);
    size_t tb = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tb, partp, kout, vin, vout, nbElem, 0, 7, s) !=
        hipSuccess)
        return MF_ERR_HIP;
    if (dev_alloc(tmp, tb) != MF_OK) return MF_ERR_NOMEM;
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, partp, kout, vin, vout, nbElem, 0, 7, s) !=
        hipSuccess)
        return MF_ERR_HIP;
    hipLaunchKernelGGL(k_color_perm, grid_for(nbElem, 256), dim3(256), 0, s, kout, vout, nbElem,
                       colorPerm, colorToElem);
    if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
    int last = 0;
    if (hipMemcpyAsync(&last, kout + nE - 1, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MF_ERR_HIP;
    *nbColors = last + 1;
    return MF_OK;
}

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

namespace {

template <typename T>
int dev_permute_rows(T *tab, const int *perm, int n, int dim, hipStream_t s)
{
    if (n < 0 || dim <= 0 || (n > 0 && (!tab || !perm))) return MF_ERR_ARG;
    if (n == 0) return MF_OK;
    DevBuf tmp, bad;
    const StreamDrain drain{s};
    if (dev_alloc(tmp, (size_t)n * dim * sizeof(T)) != MF_OK || dev_alloc(bad, sizeof(int)) != MF_OK)
        return MF_ERR_NOMEM;
    if (hipMemsetAsync(bad.p, 0, sizeof(int), s) != hipSuccess) return MF_ERR_HIP;
    hipLaunchKernelGGL(k_check_perm, grid_for(n, 256), dim3(256), 0, s, perm, n, bad.get<int>());
    int hbad = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return MF_ERR_HIP;
    if (hbad) return MF_ERR_ARG;  // tab unchanged, as the host twin leaves it
    const int64_t total = (int64_t)n * dim;
    if (hipMemcpyAsync(tmp.p, tab, (size_t)total * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return MF_ERR_HIP;
    hipLaunchKernelGGL(k_permute_rows<T>, grid_for(total, 256), dim3(256), 0, s, tmp.get<T>(), perm,
                       n, dim, tab);
    if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? MF_OK : MF_ERR_HIP;
}

}  // namespace

extern "C" int mf_dev_permute_int_2d(int *tab, const int *perm, int n, int dim, void *stream)
{
    return dev_permute_rows(tab, perm, n, dim, reinterpret_cast<hipStream_t>(stream));
}

extern "C" int mf_dev_permute_double_2d(double *tab, const int *perm, int n, int dim, void *stream)
{
    return dev_permute_rows(tab, perm, n, dim, reinterpret_cast<hipStream_t>(stream));
}
