// Host setup of the Mini-FEM step: node->element lists, the CSR pattern, the
// element->CSR-position table, greedy element coloring with its permutation, and the
// Dirichlet dof mask. These run once per mesh on the CPU, like the reference's setup
// (src/main.cc:209-351); their outputs must match the reference bit for bit because
// the CSR column order and the element order decide every value's summation order.
#include <cstdint>
#include <cstring>
#include <vector>

#include "minifem.h"
#include "abi_guard.h"

namespace {

inline bool bad_sizes(int nbElem, int nbNodes) { return nbElem < 0 || nbNodes < 0; }

}  // namespace

extern "C" int mf_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes,
                                    int *index, int *value)
{
    return mf_guarded([&]() -> int {
        if (bad_sizes(nbElem, nbNodes) || !index || (nbElem && (!elemToNode || !value)))
            return MF_ERR_ARG;
        // Two-pass counting fill: counts per node, prefix sum, then elements in ascending
        // order (the element order DC-lib's helper is taken to produce).
        std::memset(index, 0, sizeof(int) * (size_t)(nbNodes + 1));
        const int64_t nv = (int64_t)nbElem * 4;
        for (int64_t q = 0; q < nv; q++) {
            const int node = elemToNode[q] - 1;
            if (node < 0 || node >= nbNodes) return MF_ERR_ARG;
            index[node + 1]++;
        }
        for (int i = 0; i < nbNodes; i++) index[i + 1] += index[i];
        std::vector<int> cursor(index, index + nbNodes);
        for (int e = 0; e < nbElem; e++) {
            const int *v = elemToNode + (int64_t)e * 4;
            for (int k = 0; k < 4; k++) value[cursor[v[k] - 1]++] = e;
        }
        return MF_OK;
    });
}

extern "C" int mf_create_nodeToNode(int *row, int *col, const int *n2eIndex,
                                    const int *n2eValue, const int *elemToNode,
                                    int nbNodes, int nbEdges)
{
    return mf_guarded([&]() -> int {
        if (nbNodes < 0 || nbEdges < 0 || !row || !n2eIndex) return MF_ERR_ARG;
        // stamp[v] == i + 1 marks column v (1-based) as already present in row i: the
        // first-seen order of the reference's dedup scan (src/matrix.cc:67-86) without its
        // quadratic search.
        std::vector<int> stamp((size_t)nbNodes + 1, 0);
        int ctr = 0;
        for (int i = 0; i < nbNodes; i++) {
            row[i] = ctr;
            for (int j = n2eIndex[i]; j < n2eIndex[i + 1]; j++) {
                const int *v = elemToNode + (int64_t)n2eValue[j] * 4;
                for (int k = 0; k < 4; k++) {
                    const int c = v[k];
                    if (c < 1 || c > nbNodes) return MF_ERR_ARG;
                    if (stamp[c] == i + 1) continue;
                    stamp[c] = i + 1;
                    if (ctr >= nbEdges) return MF_ERR_NNZ;
                    col[ctr++] = c;
                }
            }
        }
        row[nbNodes] = ctr;
        return ctr;
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
        std::vector<int> idx((size_t)nbNodes + 1), val((size_t)nbElem * 4 + 1);
        int st = mf_create_nodeToElem(elemToNode, nbElem, nbNodes, idx.data(), val.data());
        if (st != MF_OK) return st;
        // Color bits of already-colored elements, 128 bits as two words. Greedy first fit
        // in element order (src/coloring.cc:52-76): only lower-index neighbours carry bits.
        std::vector<uint64_t> lo((size_t)nbElem, 0), hi((size_t)nbElem, 0);
        std::vector<int> part((size_t)nbElem);
        int maxColor = -1;
        for (int e = 0; e < nbElem; e++) {
            uint64_t mlo = 0, mhi = 0;
            const int *v = elemToNode + (int64_t)e * 4;
            for (int k = 0; k < 4; k++) {
                const int node = v[k] - 1;
                for (int j = idx[node]; j < idx[node + 1]; j++) {
                    const int f = val[j];
                    if (f >= e) break;  // lists are ascending: the rest are uncolored
                    mlo |= lo[f];
                    mhi |= hi[f];
                }
            }
            int color;
            if (~mlo) color = __builtin_ctzll(~mlo);
            else if (~mhi) color = 64 + __builtin_ctzll(~mhi);
            else return MF_ERR_COLORS;
            if (color >= MF_MAX_COLOR) return MF_ERR_COLORS;
            if (color < 64) lo[e] = 1ull << color;
            else hi[e] = 1ull << (color - 64);
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
