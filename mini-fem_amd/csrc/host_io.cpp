// Mesh input files and reference-norm files in the reference's formats:
// native-endian binary input (src/IO.cc:61-127) and the two-line text norm file
// (src/IO.cc:26-58), plus the synthetic jittered-cube generator and its coefficient
// fields.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "minifem.h"
#include "abi_guard.h"

namespace {

struct File {
    FILE *f;
    explicit File(FILE *fp) : f(fp) {}
    ~File() { if (f) std::fclose(f); }
};

template <typename T>
bool rd(FILE *f, T *p, int64_t n)
{
    return n == 0 || (p && std::fread(p, sizeof(T), (size_t)n, f) == (size_t)n);
}

template <typename T>
bool wr(FILE *f, const T *p, int64_t n)
{
    return n == 0 || (p && std::fwrite(p, sizeof(T), (size_t)n, f) == (size_t)n);
}

}  // namespace

extern "C" int mf_read_header(const char *path, int header[6])
{
    if (!path || !header) return MF_ERR_ARG;
    File in(std::fopen(path, "rb"));
    if (!in.f) return MF_ERR_IO;
    if (!rd(in.f, header, 6)) return MF_ERR_IO;
    for (int i = 0; i < 6; i++)
        if (header[i] < 0) return MF_ERR_IO;
    return MF_OK;
}

extern "C" int mf_read_input_data(const char *path, double *coord, int *elemToNode,
                                  int *neighborsList, int *intfIndex, int *intfNodes,
                                  int *boundNodesCode)
{
    if (!path) return MF_ERR_ARG;
    File in(std::fopen(path, "rb"));
    if (!in.f) return MF_ERR_IO;
    int h[6];
    if (!rd(in.f, h, 6)) return MF_ERR_IO;
    const int64_t nE = h[0], nN = h[1], nI = h[3], nIN = h[4];
    if (nE < 0 || nN < 0 || nI < 0 || nIN < 0) return MF_ERR_IO;
    const int64_t nl = (nI > 1 ? nI : 1) * 3;
    bool ok = rd(in.f, coord, nN * 3) && rd(in.f, elemToNode, nE * 4) &&
              rd(in.f, neighborsList, nl) && rd(in.f, intfIndex, nI + 1) &&
              rd(in.f, intfNodes, nIN) && rd(in.f, boundNodesCode, nN);
    return ok ? MF_OK : MF_ERR_IO;
}

extern "C" int mf_store_input_data(const char *path, const int header[6],
                                   const double *coord, const int *elemToNode,
                                   const int *neighborsList, const int *intfIndex,
                                   const int *intfNodes, const int *boundNodesCode)
{
    return mf_guarded([&]() -> int {
        if (!path || !header) return MF_ERR_ARG;
        for (int i = 0; i < 6; i++)
            if (header[i] < 0) return MF_ERR_ARG;
        File out(std::fopen(path, "wb"));
        if (!out.f) return MF_ERR_IO;
        const int64_t nE = header[0], nN = header[1], nI = header[3], nIN = header[4];
        const int64_t nl = (nI > 1 ? nI : 1) * 3;
        // A single rank has no interface: the reference still stores max(nbIntf,1)*3
        // neighbour ints and nbIntf+1 index ints (src/IO.cc:122-123).
        std::vector<int> zeros((size_t)(nl > nI + 1 ? nl : nI + 1), 0);
        bool ok = wr(out.f, header, 6) && wr(out.f, coord, nN * 3) &&
                  wr(out.f, elemToNode, nE * 4) &&
                  wr(out.f, neighborsList ? neighborsList : zeros.data(), nl) &&
                  wr(out.f, intfIndex ? intfIndex : zeros.data(), nI + 1) &&
                  wr(out.f, intfNodes, nIN) && wr(out.f, boundNodesCode, nN);
        return ok ? MF_OK : MF_ERR_IO;
    });
}

extern "C" int mf_read_ref_assembly(const char *path, double *matrixNorm, double *precNorm)
{
    if (!path || !matrixNorm || !precNorm) return MF_ERR_ARG;
    File in(std::fopen(path, "r"));
    if (!in.f) return MF_ERR_IO;
    return std::fscanf(in.f, "%lf %lf", matrixNorm, precNorm) == 2 ? MF_OK : MF_ERR_IO;
}

extern "C" int mf_store_ref_assembly(const char *path, double matrixNorm, double precNorm)
{
    if (!path) return MF_ERR_ARG;
    File out(std::fopen(path, "w"));
    if (!out.f) return MF_ERR_IO;
    // setprecision(17) of a default-format stream is %.17g.
    return std::fprintf(out.f, "%.17g\n%.17g\n", matrixNorm, precNorm) > 0 ? MF_OK
                                                                          : MF_ERR_IO;
}

// ---------------------------------------------------------------------------------
// Synthetic meshes. RNG: splitmix64 seeding xoshiro256**, uniform doubles from the top
// 53 bits, Box-Muller normals; no std::*_distribution (not portable across libraries).
// ---------------------------------------------------------------------------------
namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    // Integer in [0, bound) by multiply-shift.
    uint64_t below(uint64_t bound)
    {
        return (uint64_t)(((unsigned __int128)next() * bound) >> 64);
    }
    double normal()
    {
        const double u1 = 1.0 - uniform();  // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Kuhn split of a hex along (0,0,0)->(1,1,1): one tet per ordering of the axes, vertex
// list [c000, c000+e_a, c000+e_a+e_b, c111].
const int kAxisOrder[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

}  // namespace

extern "C" int mf_cube_sizes(int n, int *nbElem, int *nbNodes, int64_t *nnz)
{
    if (n < 2 || n > 1290) return MF_ERR_ARG;  // 6(n-1)^3 must fit an int
    const int64_t m = n - 1, nn = n;
    if (nbElem) *nbElem = (int)(6 * m * m * m);
    if (nbNodes) *nbNodes = (int)(nn * nn * nn);
    if (nnz) *nnz = nn * nn * nn + 2 * (3 * m * nn * nn + 3 * nn * m * m + m * m * m);
    return MF_OK;
}

extern "C" int mf_generate_cube(int n, double jitter, uint64_t geomSeed, uint64_t shuffleSeed,
                                double *coord, int *elemToNode, int *boundNodesCode)
{
    return mf_guarded([&]() -> int {
        int nE, nN;
        if (mf_cube_sizes(n, &nE, &nN, nullptr) != MF_OK || !coord || !elemToNode ||
            !boundNodesCode || jitter < 0 || jitter >= 0.5)
            return MF_ERR_ARG;
        const double h = 1.0 / (n - 1);
        Rng geo(geomSeed);
        auto id = [n](int i, int j, int k) { return i + n * (j + n * k); };
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const int v = id(i, j, k);
                    double *c = coord + (int64_t)v * 3;
                    c[0] = i * h;
                    c[1] = j * h;
                    c[2] = k * h;
                    const bool interior = i > 0 && j > 0 && k > 0 && i < n - 1 && j < n - 1 &&
                                          k < n - 1;
                    if (interior)
                        for (int d = 0; d < 3; d++) c[d] += (2.0 * geo.uniform() - 1.0) * jitter * h;
                    int code = 0;
                    if (i == n - 1 || j == n - 1 || k == n - 1) code = 10;
                    else if (i == 0) code = 52;
                    else if (j == 0) code = 53;
                    else if (k == 0) code = 54;
                    boundNodesCode[v] = code;
                }
        int64_t e = 0;
        for (int k = 0; k < n - 1; k++)
            for (int j = 0; j < n - 1; j++)
                for (int i = 0; i < n - 1; i++)
                    for (int t = 0; t < 6; t++) {
                        int p[3] = {i, j, k};
                        int *out = elemToNode + e * 4;
                        out[0] = id(p[0], p[1], p[2]) + 1;
                        p[kAxisOrder[t][0]]++;
                        out[1] = id(p[0], p[1], p[2]) + 1;
                        p[kAxisOrder[t][1]]++;
                        out[2] = id(p[0], p[1], p[2]) + 1;
                        out[3] = id(i + 1, j + 1, k + 1) + 1;
                        e++;
                    }
        if (shuffleSeed != 0) {
            // Fisher-Yates: newId[old] is a seeded random permutation of the node ids.
            std::vector<int> newId((size_t)nN);
            for (int v = 0; v < nN; v++) newId[v] = v;
            Rng sh(shuffleSeed);
            for (int v = nN - 1; v > 0; v--) {
                const int u = (int)sh.below((uint64_t)v + 1);
                std::swap(newId[v], newId[u]);
            }
            std::vector<double> c2((size_t)nN * 3);
            std::vector<int> b2((size_t)nN);
            for (int v = 0; v < nN; v++) {
                std::memcpy(&c2[(size_t)newId[v] * 3], coord + (int64_t)v * 3, 3 * sizeof(double));
                b2[newId[v]] = boundNodesCode[v];
            }
            std::memcpy(coord, c2.data(), c2.size() * sizeof(double));
            std::memcpy(boundNodesCode, b2.data(), b2.size() * sizeof(int));
            for (int64_t q = 0; q < (int64_t)nE * 4; q++) elemToNode[q] = newId[elemToNode[q] - 1] + 1;
        }
        return MF_OK;
    });
}

extern "C" int mf_generate_coef(int operatorID, int nbElem, uint64_t seed, double *coef)
{
    if ((operatorID != MF_OP_LAP && operatorID != MF_OP_ELA) || nbElem < 0 ||
        (nbElem && !coef))
        return MF_ERR_ARG;
    Rng r(seed);
    for (int e = 0; e < nbElem; e++) {
        if (operatorID == MF_OP_LAP) {
            coef[e] = std::exp(0.5 * r.normal());
        } else {
            const double E = std::exp(0.5 * r.normal());
            const double nu = 0.2 + 0.15 * r.uniform();
            coef[2 * (int64_t)e] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
            coef[2 * (int64_t)e + 1] = E / (2.0 * (1.0 + nu));
        }
    }
    return MF_OK;
}
