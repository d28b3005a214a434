// Host setup of the Mini-FEM step: node->element lists, the CSR pattern, the
// element->CSR-position table, greedy element coloring with its permutation, and the
// Dirichlet dof mask. These run once per mesh on the CPU, like the reference's setup
// (src/main.cc:209-351); their outputs must match the reference bit for bit because
// the CSR column order and the element order decide every value's summation order.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <vector>

#include "minifem.h"
#include "abi_guard.h"
#include "host_threads.h"

using mf::fork_join;
using mf::host_threads;
using mf::part_lo;

namespace {

inline bool bad_sizes(int nbElem, int nbNodes) { return nbElem < 0 || nbNodes < 0; }

}  // namespace

extern "C" int mf_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes,
                                    int *index, int *value)
{
    return mf_guarded([&]() -> int {
        if (bad_sizes(nbElem, nbNodes) || !index || (nbElem && (!elemToNode || !value)))
            return MF_ERR_ARG;
        // Counting fill with each node's elements in ascending order (the element order
        // DC-lib's helper is taken to produce). The elements are cut into T contiguous
        // chunks: each thread counts its chunk per node, the counts give every chunk its
        // place in each node's list (chunk t after chunks 0..t-1), then each thread writes
        // its own elements there -- the same lists as one pass in element order.
        const int T = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), nbElem / 65536));
        std::vector<std::vector<int>> cnt((size_t)T);
        std::atomic<int> bad{0};
        fork_join(T, [&](int t) {
            cnt[(size_t)t].assign((size_t)nbNodes, 0);
            int *c = cnt[(size_t)t].data();
            for (int64_t q = part_lo(nbElem, T, t) * 4; q < part_lo(nbElem, T, t + 1) * 4; q++) {
                const int node = elemToNode[q] - 1;
                if (node < 0 || node >= nbNodes) {
                    bad = 1;
                    return;
                }
                c[node]++;
            }
        });
        if (bad) return MF_ERR_ARG;
        // per node: the total (index) and each chunk's start (replacing its count)
        fork_join(T, [&](int t) {
            for (int64_t i = part_lo(nbNodes, T, t); i < part_lo(nbNodes, T, t + 1); i++) {
                int s = 0;
                for (int u = 0; u < T; u++) {
                    const int c = cnt[(size_t)u][(size_t)i];
                    cnt[(size_t)u][(size_t)i] = s;
                    s += c;
                }
                index[i + 1] = s;
            }
        });
        index[0] = 0;
        for (int i = 0; i < nbNodes; i++) index[i + 1] += index[i];
        fork_join(T, [&](int t) {
            int *c = cnt[(size_t)t].data();
            for (int64_t e = part_lo(nbElem, T, t); e < part_lo(nbElem, T, t + 1); e++) {
                const int *v = elemToNode + e * 4;
                for (int k = 0; k < 4; k++) {
                    const int node = v[k] - 1;
                    value[index[node] + c[node]++] = (int)e;
                }
            }
        });
        return MF_OK;
    });
}

extern "C" int mf_create_nodeToNode(int *row, int *col, const int *n2eIndex,
                                    const int *n2eValue, const int *elemToNode,
                                    int nbNodes, int nbEdges)
{
    return mf_guarded([&]() -> int {
        if (nbNodes < 0 || nbEdges < 0 || !row || !n2eIndex) return MF_ERR_ARG;
        // Rows are independent: each thread takes a contiguous range of rows, first their
        // lengths, then (after a prefix sum) their columns. stamp[v] == i + 1 marks column
        // v (1-based) as already present in row i: the first-seen order of the
        // reference's dedup scan (src/matrix.cc:67-86) without its quadratic search.
        const int T = std::max(1, std::min(host_threads(), nbNodes / 16384));
        std::atomic<int> bad{0};
        auto rows = [&](int t, bool fill) {
            std::vector<int> stamp((size_t)nbNodes + 1, 0);
            for (int i = (int)part_lo(nbNodes, T, t); i < (int)part_lo(nbNodes, T, t + 1); i++) {
                int ctr = fill ? row[i] : 0;
                for (int j = n2eIndex[i]; j < n2eIndex[i + 1]; j++) {
                    const int *v = elemToNode + (int64_t)n2eValue[j] * 4;
                    for (int k = 0; k < 4; k++) {
                        const int c = v[k];
                        if (c < 1 || c > nbNodes) {
                            bad = 1;
                            return;
                        }
                        if (stamp[c] == i + 1) continue;
                        stamp[c] = i + 1;
                        if (fill) col[ctr] = c;
                        ctr++;
                    }
                }
                if (!fill) row[i + 1] = ctr;  // the row's length, summed below
            }
        };
        fork_join(T, [&](int t) { rows(t, false); });
        if (bad) return MF_ERR_ARG;
        row[0] = 0;
        int64_t total = 0;
        for (int i = 0; i < nbNodes; i++) {
            total += row[i + 1];
            if (total > nbEdges) return MF_ERR_NNZ;  // the pattern does not fit col[]
            row[i + 1] = (int)total;
        }
        if (total > 0 && !col) return MF_ERR_ARG;
        fork_join(T, [&](int t) { rows(t, true); });
        return (int)total;
    });
}

extern "C" int mf_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                                    int *elemToEdge, int nbElem)
{
    if (nbElem < 0 || (nbElem && (!row || !col || !elemToNode || !elemToEdge)))
        return MF_ERR_ARG;
    for (int e = 0; e < nbElem; e++) {
        const int *v = elemToNode + (int64_t)e * 4;
        int *out = elemToEdge + (int64_t)e * 16;
        for (int j = 0; j < 4; j++) {
            const int r = v[j] - 1;
            for (int k = 0; k < 4; k++) {
                int pos = -1;
                for (int l = row[r]; l < row[r + 1]; l++)
                    if (col[l] == v[k]) { pos = l; break; }
                if (pos < 0) return MF_ERR_ARG;  // pattern does not hold this element
                out[j * 4 + k] = pos;
            }
        }
    }
    return MF_OK;
}

extern "C" int mf_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                                    int nbElem, int nbNodes, int *colorToElem,
                                    int *nbColors)
{
    return mf_guarded([&]() -> int {
        if (bad_sizes(nbElem, nbNodes) || !colorPerm || !colorToElem || !nbColors)
            return MF_ERR_ARG;
        // Greedy first fit in element order (src/coloring.cc:46-81): element e takes the
        // lowest color that no element sharing a node with it holds. The reference ORs its
        // neighbours' color bits, and those not colored yet hold none (src/coloring.cc:48,
        // :57-58), so only lower-index neighbours count. Their bits, gathered per node, are
        // the OR of the colors every earlier element at that node took: a 128-bit mask per
        // node (two words), updated as each element is colored. Same colors, O(1) work per
        // element instead of a walk over its neighbours.
        std::vector<uint64_t> lo((size_t)nbNodes, 0), hi((size_t)nbNodes, 0);
        std::vector<int> part((size_t)nbElem);
        int maxColor = -1;
        for (int e = 0; e < nbElem; e++) {
            const int *v = elemToNode + (int64_t)e * 4;
            uint64_t mlo = 0, mhi = 0;
            for (int k = 0; k < 4; k++) {
                const int node = v[k] - 1;
                if (node < 0 || node >= nbNodes) return MF_ERR_ARG;
                mlo |= lo[node];
                mhi |= hi[node];
            }
            int color;
            if (~mlo) color = __builtin_ctzll(~mlo);
            else if (~mhi) color = 64 + __builtin_ctzll(~mhi);
            else return MF_ERR_COLORS;
            if (color >= MF_MAX_COLOR) return MF_ERR_COLORS;
            const uint64_t blo = color < 64 ? 1ull << color : 0, bhi = color < 64 ? 0 : 1ull << (color - 64);
            for (int k = 0; k < 4; k++) {
                lo[v[k] - 1] |= blo;
                hi[v[k] - 1] |= bhi;
            }
            part[e] = color;
            if (color > maxColor) maxColor = color;
        }
        const int nc = maxColor + 1;
        std::vector<int> count((size_t)(nc > 0 ? nc : 1), 0);
        for (int e = 0; e < nbElem; e++) count[part[e]]++;
        colorToElem[0] = 0;
        for (int c = 0; c < nc; c++) colorToElem[c + 1] = colorToElem[c] + count[c];
        std::vector<int> next(colorToElem, colorToElem + (nc > 0 ? nc : 1));
        for (int e = 0; e < nbElem; e++) colorPerm[e] = next[part[e]]++;
        if (colorPart) std::memcpy(colorPart, part.data(), sizeof(int) * (size_t)nbElem);
        *nbColors = nc;
        return MF_OK;
    });
}

template <typename T>
static int permute_rows(T *tab, const int *perm, int n, int dim)
{
    if (n < 0 || dim <= 0 || (n && (!tab || !perm))) return MF_ERR_ARG;
    std::vector<T> tmp((size_t)n * dim);
    for (int i = 0; i < n; i++) {
        const int p = perm[i];
        if (p < 0 || p >= n) return MF_ERR_ARG;
        std::memcpy(&tmp[(size_t)p * dim], tab + (size_t)i * dim, sizeof(T) * dim);
    }
    std::memcpy(tab, tmp.data(), sizeof(T) * tmp.size());
    return MF_OK;
}

extern "C" int mf_permute_int_2d(int *tab, const int *perm, int n, int dim)
{
    return mf_guarded([&]() -> int {
        return permute_rows(tab, perm, n, dim);
    });
}

extern "C" int mf_permute_double_2d(double *tab, const int *perm, int n, int dim)
{
    return mf_guarded([&]() -> int {
        return permute_rows(tab, perm, n, dim);
    });
}

extern "C" int mf_boundary_mask(const int *boundNodesCode, int nbNodes, int nbBoundNodes,
                                int *checkBounds)
{
    if (nbNodes < 0 || nbBoundNodes < 0 || (nbNodes && (!boundNodesCode || !checkBounds)))
        return MF_ERR_ARG;
    std::memset(checkBounds, 0, sizeof(int) * (size_t)nbNodes * 3);
    // DQMRD4 lists at most nbBoundNodes coded nodes in ascending order
    // (src/Fortran/qdmrd4.f:8-16); E_ESSBCM masks dofs by code
    // (src/Fortran/e_cgmelissa.F:25-38; codes src/Fortran/DefMesh_jlog.h:10-20).
    int listed = 0;
    for (int i = 0; i < nbNodes && listed < nbBoundNodes; i++) {
        const int code = boundNodesCode[i];
        if (code == 0) continue;
        listed++;
        const int64_t N = nbNodes;
        if (code == 52) checkBounds[0 * N + i] = 1;
        else if (code == 53) checkBounds[1 * N + i] = 1;
        else if (code == 54) checkBounds[2 * N + i] = 1;
        else {
            checkBounds[0 * N + i] = 1;
            checkBounds[1 * N + i] = 1;
            checkBounds[2 * N + i] = 1;
        }
    }
    return MF_OK;
}

extern "C" double mf_compute_double_norm(const double *tab, int64_t size)
{
    double s = 0;
    for (int64_t i = 0; i < size; i++) s += tab[i] * tab[i];
    return __builtin_sqrt(s);
}
