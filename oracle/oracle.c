/*
 * CPU oracle: plain-C restatement of Mini-FEM's assembly + block-Jacobi step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -O2 -ffp-contract=off so
 * that every multiply and add rounds separately, in the order the reference source
 * writes them; the GPU kernels are compiled the same way, which makes the GPU
 * result comparable bit for bit.
 *
 * Pinning of each function (what checks it besides this file):
 *   or_elem_coef, or_assembly_*    closed-form unit-tet KAT, row-sum/symmetry/scaling
 *                                  invariants, brute-force dense assembly on tiny meshes
 *                                  (tests/test_oracle_kat.py)
 *   or_create_nodeToElem/ToNode    two-tet KAT, closed-form nnz of the Kuhn cube
 *   or_coloring_creation           validity + first-fit property + two-tet KAT
 *   or_create_elemToEdge           col[pos] == node check for every (elem, j, k)
 *   or_prec_init / or_prec_inversion  lap reciprocal KAT, closed-form 3x3 inverse,
 *                                  A*inv(A) = I, masked unit rows/cols
 *   or_dqmrd4 / or_e_essbcm        code-table KAT (52/53/54/10/0)
 *   or_compute_double_norm         brute-force numpy sum of squares
 *   or_assembly_mixed              parity unpinned bitwise against the reference
 *                                  (the reference has no mixed form); pinned to the
 *                                  FP64 oracle within 1e-6 relative Frobenius error.
 *   DGETRF/DGETRI restatement      parity unpinned bitwise against MKL (unavailable);
 *                                  pinned by the closed form and A*inv(A) = I.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <x86intrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* setup                                                                      */
/* ------------------------------------------------------------------------- */

/* DC-lib's DC_create_nodeToElem (called at src/main.cc:247 and
 * src/coloring.cc:90) is not in the reference tree. Adopted semantics: for each
 * 0-based node, the 0-based ids of the elements that contain it, in ascending element
 * order (a two-pass counting fill over elements in order, vertices in local order). */
int or_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes,
                         int *n2eIndex, int *n2eValue)
{
    for (int i = 0; i <= nbNodes; i++) n2eIndex[i] = 0;
    for (int64_t e = 0; e < (int64_t)nbElem * OR_DIM_ELEM; e++) {
        int node = elemToNode[e] - 1;
        if (node < 0 || node >= nbNodes) return OR_ERR_ARG;
        n2eIndex[node + 1]++;
    }
    for (int i = 0; i < nbNodes; i++) n2eIndex[i + 1] += n2eIndex[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)(nbNodes > 0 ? nbNodes : 1));
    if (!fill) return OR_ERR_ARG;
    memcpy(fill, n2eIndex, sizeof(int) * (size_t)nbNodes);
    for (int e = 0; e < nbElem; e++)
        for (int k = 0; k < OR_DIM_ELEM; k++) {
            int node = elemToNode[e * OR_DIM_ELEM + k] - 1;
            n2eValue[fill[node]++] = e;
        }
    free(fill);
    return OR_OK;
}

/* src/matrix.cc:55-91 (create_nodeToNode): node->node adjacency in first-seen order
 * over the incident elements' vertices, diagonal included, 1-based column ids.
 * The reference writes into a buffer sized by the header's nbEdges without a bounds
 * check (src/main.cc:246, src/matrix.cc:81); here an overflow returns OR_ERR_NNZ.
 * Returns the number of entries written (>= 0) or an error (< 0). */
int or_create_nodeToNode(int *row, int *col, const int *n2eIndex, const int *n2eValue,
                         const int *elemToNode, int nbNodes, int nbEdgesCap)
{
    int ctr = 0;
    for (int i = 0; i < nbNodes; i++) {
        int nbMax = (n2eIndex[i + 1] - n2eIndex[i]) * OR_DIM_ELEM;
        int *nb = (int *)malloc(sizeof(int) * (size_t)(nbMax > 0 ? nbMax : 1));
        int n = 0;
        row[i] = ctr;
        for (int j = n2eIndex[i]; j < n2eIndex[i + 1]; j++) {
            int e = n2eValue[j];
            for (int k = 0; k < OR_DIM_ELEM; k++) {
                int v = elemToNode[e * OR_DIM_ELEM + k];
                int isNew = 1;
                for (int l = 0; l < n; l++)
                    if (v == nb[l]) isNew = 0;
                if (isNew) {
                    if (ctr >= nbEdgesCap) { free(nb); return OR_ERR_NNZ; }
                    col[ctr++] = v;
                    nb[n++] = v;
                }
            }
        }
        free(nb);
    }
    row[nbNodes] = ctr;
    return ctr;
}

/* src/matrix.cc:25-52 (create_elemToEdge): for each element the 16 CSR positions of
 * its (j,k) vertex pairs, row-major in (j,k). */
void or_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                          int *elemToEdge, int nbElem)
{
    for (int i = 0; i < nbElem; i++) {
        int ctr = 0;
        for (int j = 0; j < OR_DIM_ELEM; j++) {
            int n1 = elemToNode[i * OR_DIM_ELEM + j] - 1;
            for (int k = 0; k < OR_DIM_ELEM; k++) {
                int n2 = elemToNode[i * OR_DIM_ELEM + k] - 1;
                for (int l = row[n1]; l < row[n1 + 1]; l++) {
                    if (col[l] == n2 + 1) {
                        elemToEdge[(int64_t)i * OR_VALUES_PER_ELEM + ctr] = l;
                        ctr++;
                        break;
                    }
                }
            }
        }
    }
}

/* src/coloring.cc:84-109 (coloring_creation), :46-81 (create_longest_color_part),
 * :27-42 (fill_color_index).
 *   - neighbour relation (DC_create_elemToElem, external): elements sharing >= 1 node.
 *     First-fit only reads colors of already-colored (lower-index) elements, and the
 *     128-bit OR is insensitive to list order, duplicates and self-inclusion, so the
 *     result does not depend on how DC-lib orders or dedups its lists.
 *   - color = lowest zero bit of the OR of neighbour color bits; >= 128 is an error.
 *   - colorToElem[c] = first (permuted) element of color c, colorToElem[nbColors]=nbElem.
 *   - colorPerm (DC_create_permutation, external): stable counting sort by color,
 *     stored old -> new (colorPerm[old] = new position).
 * colorPart may be NULL. */
int or_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                         int nbElem, int nbNodes, int *colorToElem, int *nbColors)
{
    int *idx = (int *)malloc(sizeof(int) * (size_t)(nbNodes + 1));
    int *val = (int *)malloc(sizeof(int) * (size_t)nbElem * OR_DIM_ELEM + 1);
    unsigned __int128 *bits = (unsigned __int128 *)calloc((size_t)nbElem + 1, 16);
    int *part = colorPart ? colorPart : (int *)malloc(sizeof(int) * (size_t)(nbElem + 1));
    int status = OR_OK;
    if (!idx || !val || !bits || !part) { status = OR_ERR_ARG; goto done; }
    status = or_create_nodeToElem(elemToNode, nbElem, nbNodes, idx, val);
    if (status != OR_OK) goto done;

    int maxColor = 0;
    for (int i = 0; i < nbElem; i++) {
        unsigned __int128 mask = 1, nc = 0;
        for (int k = 0; k < OR_DIM_ELEM; k++) {
            int node = elemToNode[i * OR_DIM_ELEM + k] - 1;
            for (int j = idx[node]; j < idx[node + 1]; j++) nc |= bits[val[j]];
        }
        int color = 0;
        while (nc & mask) { nc >>= 1; color++; }
        if (color >= OR_MAX_COLOR) { status = OR_ERR_COLORS; goto done; }
        bits[i] = mask << color;
        part[i] = color;
        if (color > maxColor) maxColor = color;
    }
    *nbColors = nbElem > 0 ? maxColor + 1 : 0;

    /* fill_color_index with offset 0 */
    {
        int count[OR_MAX_COLOR] = {0};
        for (int i = 0; i < nbElem; i++) count[part[i]]++;
        colorToElem[0] = 0;
        for (int c = 1; c <= *nbColors; c++) colorToElem[c] = colorToElem[c - 1] + count[c - 1];
        int next[OR_MAX_COLOR];
        for (int c = 0; c < *nbColors; c++) next[c] = colorToElem[c];
        for (int i = 0; i < nbElem; i++) colorPerm[i] = next[part[i]]++;
    }
done:
    free(idx); free(val); free(bits);
    if (!colorPart) free(part);
    return status;
}

/* DC_permute_int_2d_array as applied at src/main.cc:229: row `old` of tab moves to
 * row perm[old]. */
void or_permute_int_2d(int *tab, const int *perm, int n, int dim)
{
    int *tmp = (int *)malloc(sizeof(int) * (size_t)n * dim + 1);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) tmp[(int64_t)perm[i] * dim + d] = tab[(int64_t)i * dim + d];
    memcpy(tab, tmp, sizeof(int) * (size_t)n * dim);
    free(tmp);
}

void or_permute_double_2d(double *tab, const int *perm, int n, int dim)
{
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n * dim + 1);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) tmp[(int64_t)perm[i] * dim + d] = tab[(int64_t)i * dim + d];
    memcpy(tab, tmp, sizeof(double) * (size_t)n * dim);
    free(tmp);
}

/* src/Fortran/qdmrd4.f:1-23 (DQMRD4): list (1-based) the nodes whose code != 0.
 * *nbBound is the capacity on entry and the count on exit; returns -1 when the
 * count exceeds the capacity (only the first `capacity` ids are stored). */
int or_dqmrd4(int nbNodes, const int *code, int *nbBound, int *boundList)
{
    int cap = *nbBound, n = 0;
    for (int i = 1; i <= nbNodes; i++) {
        if (code[i - 1] != 0) {
            n++;
            if (n <= cap) boundList[n - 1] = i;
        }
    }
    *nbBound = n;
    return n > cap ? -1 : 0;
}

/* src/Fortran/e_cgmelissa.F:1-51 (E_ESSBCM, NDIM=3 branch) with the codes of
 * src/Fortran/DefMesh_jlog.h:10-20: 52 masks dof x, 53 dof y, 54 dof z, any other
 * nonzero code masks all three. markf(NBP,3) -> checkBounds[k*N + i]. */
void or_e_essbcm(int nbNodes, int nbBound, const int *boundList, const int *code,
                 int *checkBounds)
{
    for (int64_t i = 0; i < (int64_t)nbNodes * 3; i++) checkBounds[i] = 0;
    for (int b = 0; b < nbBound; b++) {
        int nl = boundList[b] - 1;
        int j = code[nl];
        if (j == 52) checkBounds[0 * (int64_t)nbNodes + nl] = 1;
        else if (j == 53) checkBounds[1 * (int64_t)nbNodes + nl] = 1;
        else if (j == 54) checkBounds[2 * (int64_t)nbNodes + nl] = 1;
        else {
            checkBounds[0 * (int64_t)nbNodes + nl] = 1;
            checkBounds[1 * (int64_t)nbNodes + nl] = 1;
            checkBounds[2 * (int64_t)nbNodes + nl] = 1;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* element integration and assembly                                          */
/* ------------------------------------------------------------------------- */

/* src/assembly.cc:85-121 (elem_coef_seq): a=v0-v3, b=v2-v3, c=v1-v3; g0=b x c,
 * g1=a x b, g2=c x a, g3=-(g0+g1+g2); det = a.g0; every g multiplied by (1./det). */
void or_elem_coef(double g[4][3], const double *coord, const int *elemToNode, int elem)
{
    double c[4][3];
    for (int i = 0; i < 4; i++) {
        int node = elemToNode[(int64_t)elem * 4 + i] - 1;
        for (int j = 0; j < 3; j++) c[i][j] = coord[(int64_t)node * 3 + j];
    }
    double xa = c[0][0] - c[3][0], xb = c[2][0] - c[3][0], xc = c[1][0] - c[3][0];
    double ya = c[0][1] - c[3][1], yb = c[2][1] - c[3][1], yc = c[1][1] - c[3][1];
    double za = c[0][2] - c[3][2], zb = c[2][2] - c[3][2], zc = c[1][2] - c[3][2];
    g[0][0] = yb * zc - yc * zb;
    g[0][1] = zb * xc - zc * xb;
    g[0][2] = xb * yc - xc * yb;
    g[1][0] = ya * zb - yb * za;
    g[1][1] = za * xb - zb * xa;
    g[1][2] = xa * yb - xb * ya;
    g[2][0] = yc * za - ya * zc;
    g[2][1] = zc * xa - za * xc;
    g[2][2] = xc * ya - xa * yc;
    for (int d = 0; d < 3; d++) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
    double vol = xa * g[0][0] + ya * g[0][1] + za * g[0][2];
    double r = 1. / vol;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++) g[i][j] *= r;
}

/* CSR position of (node1 row, node2 column): elemToEdge when given (-DOPTIMIZED,
 * src/assembly.cc:532-544) else the linear row search (src/assembly.cc:547-553). */
static inline int or_pos(const int *row, const int *col, const int *elemToEdge,
                         const int *elemToNode, int e, int j, int k)
{
    if (elemToEdge) return elemToEdge[(int64_t)e * 16 + j * 4 + k];
    int n1 = elemToNode[(int64_t)e * 4 + j] - 1;
    int n2 = elemToNode[(int64_t)e * 4 + k];
    for (int l = row[n1]; l < row[n1 + 1]; l++)
        if (col[l] == n2) return l;
    return -1;
}

/* src/assembly.cc:481-564 (assembly_lap_seq, REF/COLORING body): for each (j,k),
 * values[pos] += (gj.x*gk.x + gj.y*gk.y) + gj.z*gk.z. With a per-element coefficient
 * (extension, not in the reference) the sum is multiplied by coef[e]; coef == NULL or
 * coef[e] == 1.0 reproduces the reference bit for bit. */
void or_assembly_lap_seq(const double *coord, double *values, const int *row,
                         const int *col, const int *elemToNode, const int *elemToEdge,
                         const double *coef, int firstElem, int lastElem)
{
    for (int e = firstElem; e <= lastElem; e++) {
        double g[4][3];
        or_elem_coef(g, coord, elemToNode, e);
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) {
                int l = or_pos(row, col, elemToEdge, elemToNode, e, j, k);
                if (l < 0) continue;
                double v = g[j][0] * g[k][0] + g[j][1] * g[k][1] + g[j][2] * g[k][2];
                if (coef) v = coef[e] * v;
                values[l] += v;
            }
    }
}

/* src/assembly.cc:328-453 (assembly_ela_seq): 3x3 block per (j,k), row-major,
 *   diag (d,d): (gj.d gk.d)*(l+2m) + sum_{e!=d} (gj.e gk.e)*m   (x,y,z order)
 *   off  (d,e): (gj.d gk.e)*(l+m)
 * with (l+2m, l+m, m) = (2.25, 1.25, 1.0) in the reference (src/assembly.cc:386-409).
 * Extension: coef[2e] = lambda, coef[2e+1] = mu per element; the constants are
 * recovered bit for bit with lambda = 0.25, mu = 1.0 (all exact in binary). */
void or_assembly_ela_seq(const double *coord, double *values, const int *row,
                         const int *col, const int *elemToNode, const int *elemToEdge,
                         const double *coef, int firstElem, int lastElem)
{
    for (int e = firstElem; e <= lastElem; e++) {
        double g[4][3];
        or_elem_coef(g, coord, elemToNode, e);
        double cd = 2.25, co = 1.25, cm = 1.0;
        if (coef) {
            double lam = coef[2 * (int64_t)e], mu = coef[2 * (int64_t)e + 1];
            cd = lam + 2.0 * mu;
            co = lam + mu;
            cm = mu;
        }
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) {
                int l = or_pos(row, col, elemToEdge, elemToNode, e, j, k);
                if (l < 0) continue;
                double *v = values + (int64_t)l * 9;
                double p00 = g[j][0] * g[k][0], p11 = g[j][1] * g[k][1],
                       p22 = g[j][2] * g[k][2];
                v[0] += p00 * cd + p11 * cm + p22 * cm;
                v[1] += g[j][0] * g[k][1] * co;
                v[2] += g[j][0] * g[k][2] * co;
                v[3] += g[j][1] * g[k][0] * co;
                v[4] += p00 * cm + p11 * cd + p22 * cm;
                v[5] += g[j][1] * g[k][2] * co;
                v[6] += g[j][2] * g[k][0] * co;
                v[7] += g[j][2] * g[k][1] * co;
                v[8] += p00 * cm + p11 * cm + p22 * cd;
            }
    }
}

/* src/assembly.cc:615-659 (assembly, REF branch): zero values[0 .. nbEdges*dim), then
 * the element loop in element order. */
void or_assembly(const double *coord, double *values, const int *row, const int *col,
                 const int *elemToNode, const int *elemToEdge, const double *coef,
                 int nbElem, int nbEdges, int operatorDim, int operatorID)
{
    for (int64_t i = 0; i < (int64_t)nbEdges * operatorDim; i++) values[i] = 0;
    if (operatorID == 0)
        or_assembly_lap_seq(coord, values, row, col, elemToNode, elemToEdge, coef, 0, nbElem - 1);
    else
        or_assembly_ela_seq(coord, values, row, col, elemToNode, elemToEdge, coef, 0, nbElem - 1);
}

/* src/assembly.cc:593-611 + :660-673 (COLORING branch): parallel zero, then colors in
 * sequence; elements of a color run serially (OpenMP pragma disabled in the
 * reference, src/assembly.cc:362, :516) unless parallelElements != 0 (the labelled
 * per-color parallel variant). The result is identical either way because no two
 * elements of one color share a CSR entry. */
void or_assembly_coloring(const double *coord, double *values, const int *row,
                          const int *col, const int *elemToNode, const int *elemToEdge,
                          const double *coef, const int *colorToElem, int nbColors,
                          int nbEdges, int operatorDim, int operatorID,
                          int parallelElements)
{
    int64_t n = (int64_t)nbEdges * operatorDim;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) values[i] = 0;
    for (int c = 0; c < nbColors; c++) {
        int first = colorToElem[c], last = colorToElem[c + 1] - 1;
        if (!parallelElements) {
            if (operatorID == 0)
                or_assembly_lap_seq(coord, values, row, col, elemToNode, elemToEdge, coef, first, last);
            else
                or_assembly_ela_seq(coord, values, row, col, elemToNode, elemToEdge, coef, first, last);
        } else {
#pragma omp parallel for schedule(static, 256)
            for (int e = first; e <= last; e++) {
                if (operatorID == 0)
                    or_assembly_lap_seq(coord, values, row, col, elemToNode, elemToEdge, coef, e, e);
                else
                    or_assembly_ela_seq(coord, values, row, col, elemToNode, elemToEdge, coef, e, e);
            }
        }
    }
}

/* Mixed-precision restatement (no reference counterpart): Jacobian, determinant and
 * 1/det in FP64 exactly as or_elem_coef; the scaled gradients and the coefficients
 * are rounded to FP32; the per-entry products and the sums inside one element are FP32;
 * each entry is converted to FP64 and added into the FP64 CSR values. */
void or_assembly_mixed(const double *coord, double *values, const int *row,
                       const int *col, const int *elemToNode, const int *elemToEdge,
                       const double *coef, int nbElem, int nbEdges, int operatorDim,
                       int operatorID)
{
    for (int64_t i = 0; i < (int64_t)nbEdges * operatorDim; i++) values[i] = 0;
    for (int e = 0; e < nbElem; e++) {
        double gd[4][3];
        or_elem_coef(gd, coord, elemToNode, e);
        float g[4][3];
        for (int i = 0; i < 4; i++)
            for (int d = 0; d < 3; d++) g[i][d] = (float)gd[i][d];
        if (operatorID == 0) {
            float kf = coef ? (float)coef[e] : 1.0f;
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 4; k++) {
                    int l = or_pos(row, col, elemToEdge, elemToNode, e, j, k);
                    if (l < 0) continue;
                    float v = g[j][0] * g[k][0] + g[j][1] * g[k][1] + g[j][2] * g[k][2];
                    v = kf * v;
                    values[l] += (double)v;
                }
        } else {
            double cdd = 2.25, cod = 1.25, cmd = 1.0;
            if (coef) {
                double lam = coef[2 * (int64_t)e], mu = coef[2 * (int64_t)e + 1];
                cdd = lam + 2.0 * mu;
                cod = lam + mu;
                cmd = mu;
            }
            float cd = (float)cdd, co = (float)cod, cm = (float)cmd;
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 4; k++) {
                    int l = or_pos(row, col, elemToEdge, elemToNode, e, j, k);
                    if (l < 0) continue;
                    double *v = values + (int64_t)l * 9;
                    float p00 = g[j][0] * g[k][0], p11 = g[j][1] * g[k][1],
                          p22 = g[j][2] * g[k][2];
                    float b[9];
                    b[0] = p00 * cd + p11 * cm + p22 * cm;
                    b[1] = g[j][0] * g[k][1] * co;
                    b[2] = g[j][0] * g[k][2] * co;
                    b[3] = g[j][1] * g[k][0] * co;
                    b[4] = p00 * cm + p11 * cd + p22 * cm;
                    b[5] = g[j][1] * g[k][2] * co;
                    b[6] = g[j][2] * g[k][0] * co;
                    b[7] = g[j][2] * g[k][1] * co;
                    b[8] = p00 * cm + p11 * cm + p22 * cd;
                    for (int q = 0; q < 9; q++) v[q] += (double)b[q];
                }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* preconditioner                                                             */
/* ------------------------------------------------------------------------- */

/* src/preconditioner.cc:52-87 (prec_init): prec = 0, then for each node the d values
 * of the CSR entry whose 1-based column equals node+1. */
void or_prec_init(double *prec, const double *values, const int *row, const int *col,
                  int nbNodes, int operatorDim, int parallel)
{
    int64_t n = (int64_t)nbNodes * operatorDim;
#pragma omp parallel for schedule(static) if (parallel)
    for (int64_t i = 0; i < n; i++) prec[i] = 0;
#pragma omp parallel for schedule(static) if (parallel)
    for (int i = 0; i < nbNodes; i++) {
        for (int j = row[i]; j < row[i + 1]; j++) {
            if (col[j] - 1 == i) {
                for (int k = 0; k < operatorDim; k++)
                    prec[(int64_t)i * operatorDim + k] = values[(int64_t)j * operatorDim + k];
                break;
            }
        }
    }
}

/* Reference LAPACK DGETF2/DGETRF2 for a 3x3 column-major matrix, A(r,c) = a[c*3+r]:
 * IDAMAX pivot (first maximal |.|), row swap, multipliers scaled by the reciprocal
 * ONE/A(j,j) (DSCAL branch; the divide branch below SFMIN = DBL_MIN), trailing update
 * a(i,c) = a(i,c) + l_i*(-u_c) (DGER). Returns INFO (0 or the 1-based first zero
 * pivot). */
#define OR_A(r, c) a[(c) * 3 + (r)]
static int or_dgetrf3(double *a, int ipiv[3])
{
    int info = 0;
    for (int j = 0; j < 3; j++) {
        int jp = j;
        double amax = fabs(OR_A(j, j));
        for (int i = j + 1; i < 3; i++)
            if (fabs(OR_A(i, j)) > amax) { amax = fabs(OR_A(i, j)); jp = i; }
        ipiv[j] = jp;
        if (OR_A(jp, j) != 0.0) {
            if (jp != j)
                for (int c = 0; c < 3; c++) {
                    double t = OR_A(j, c); OR_A(j, c) = OR_A(jp, c); OR_A(jp, c) = t;
                }
            if (j < 2) {
                if (fabs(OR_A(j, j)) >= 2.2250738585072014e-308) {
                    double r = 1.0 / OR_A(j, j);
                    for (int i = j + 1; i < 3; i++) OR_A(i, j) *= r;
                } else {
                    for (int i = j + 1; i < 3; i++) OR_A(i, j) = OR_A(i, j) / OR_A(j, j);
                }
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j < 2) {
            for (int c = j + 1; c < 3; c++) {
                if (OR_A(j, c) != 0.0) {
                    double t = -OR_A(j, c);
                    for (int i = j + 1; i < 3; i++) OR_A(i, c) = OR_A(i, c) + OR_A(i, j) * t;
                }
            }
        }
    }
    return info;
}

/* Reference LAPACK DGETRI (unblocked path, taken since NB >= N) on the DGETRF output:
 * DTRTRI/DTRTI2 inverts U column by column (DTRMV with the X(J2) != 0 test, then DSCAL
 * by -inv(A(j,j))); then for j = N..1 the strictly lower part of column j is moved to
 * WORK and A(:,j) += sum_c (-WORK(c)) * A(:,c) (DGEMV, columns in order); finally the
 * column interchanges are undone in reverse order. A zero U diagonal returns INFO > 0
 * and leaves A unchanged (DTRTRI's singularity test). */
static int or_dgetri3(double *a, const int ipiv[3])
{
    for (int j = 0; j < 3; j++)
        if (OR_A(j, j) == 0.0) return j + 1;
    for (int j = 0; j < 3; j++) {
        OR_A(j, j) = 1.0 / OR_A(j, j);
        double ajj = -OR_A(j, j);
        /* x = A(0:j-1, j) := U(0:j-1,0:j-1)^-1-part * x  (DTRMV upper, no transpose) */
        for (int j2 = 0; j2 < j; j2++) {
            if (OR_A(j2, j) != 0.0) {
                double t = OR_A(j2, j);
                for (int i = 0; i < j2; i++) OR_A(i, j) = OR_A(i, j) + t * OR_A(i, j2);
                OR_A(j2, j) = OR_A(j2, j) * OR_A(j2, j2);
            }
        }
        for (int i = 0; i < j; i++) OR_A(i, j) *= ajj;
    }
    double work[3];
    for (int j = 2; j >= 0; j--) {
        for (int i = j + 1; i < 3; i++) { work[i] = OR_A(i, j); OR_A(i, j) = 0.0; }
        for (int c = j + 1; c < 3; c++) {
            double t = -work[c];
            for (int i = 0; i < 3; i++) OR_A(i, j) = OR_A(i, j) + t * OR_A(i, c);
        }
    }
    for (int j = 1; j >= 0; j--) {
        int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < 3; i++) {
                double t = OR_A(i, j); OR_A(i, j) = OR_A(i, jp); OR_A(i, jp) = t;
            }
    }
    return 0;
}
#undef OR_A

/* src/Fortran/elasclpr.f:2-56 (ela_invert_prec) for node curNode1 (1-based):
 * prec(3,3,N) column-major, i.e. prec[(i-1)*9 + kj*3 + ki] = prec(ki+1, kj+1, i).
 * Mask every dof with markf(i,k) != 0 (zero row k and column k, diagonal 1); if the
 * row holds its diagonal entry, invert the block with DGETRF + DGETRI. Returns the
 * LAPACK info (nonzero for a singular block; the reference only prints it). */
int or_ela_invert_prec(int nbNodes, const int *row, const int *col, double *prec,
                       const int *checkBounds, int curNode1)
{
    int i = curNode1 - 1;
    double *p = prec + (int64_t)i * 9;
    for (int ki = 0; ki < 3; ki++) {
        if (checkBounds[(int64_t)ki * nbNodes + i] != 0) {
            for (int kj = 0; kj < 3; kj++) {
                p[kj * 3 + ki] = 0.0;
                p[ki * 3 + kj] = 0.0;
            }
            p[ki * 3 + ki] = 1.0;
        }
    }
    int found = 0;
    for (int m = row[i]; m < row[i + 1]; m++)
        if (col[m] == curNode1) { found = 1; break; }
    if (!found) return 0;
    double a[9];
    int ipiv[3];
    memcpy(a, p, sizeof a);
    int info = or_dgetrf3(a, ipiv);
    int info2 = or_dgetri3(a, ipiv);
    memcpy(p, a, sizeof a);
    return info ? info : info2;
}

/* src/preconditioner.cc:25-49 (prec_inversion): lap prec[i] = 1.0/prec[i]; ela one
 * ela_invert_prec call per node. Returns the number of singular ela blocks. */
int or_prec_inversion(double *prec, const int *row, const int *col,
                      const int *checkBounds, int nbNodes, int operatorID, int parallel)
{
    int nsing = 0;
#pragma omp parallel for schedule(static) if (parallel) reduction(+ : nsing)
    for (int i = 0; i < nbNodes; i++) {
        if (operatorID == 0) {
            prec[i] = 1.0 / prec[i];
        } else {
            if (or_ela_invert_prec(nbNodes, row, col, prec, checkBounds, i + 1) != 0) nsing++;
        }
    }
    return nsing;
}

/* src/FEM.cc:48-56 (compute_double_norm): serial sqrt(sum pow(x,2)). */
double or_compute_double_norm(const double *tab, int64_t size)
{
    double norm = 0;
    for (int64_t i = 0; i < size; i++) norm += pow(tab[i], 2);
    return sqrt(norm);
}

/* ------------------------------------------------------------------------- */
/* mesh file (src/IO.cc:61-96)                                                */
/* ------------------------------------------------------------------------- */

int or_read_header(const char *path, int hdr[6])
{
    FILE *f = fopen(path, "rb");
    if (!f) return OR_ERR_IO;
    size_t n = fread(hdr, sizeof(int), 6, f);
    fclose(f);
    return n == 6 ? OR_OK : OR_ERR_IO;
}

int or_read_input_data(const char *path, double *coord, int *elemToNode,
                       int *neighborsList, int *intfIndex, int *intfNodes,
                       int *boundNodesCode)
{
    int h[6];
    FILE *f = fopen(path, "rb");
    if (!f) return OR_ERR_IO;
    int ok = fread(h, sizeof(int), 6, f) == 6;
    int nE = h[0], nN = h[1], nI = h[3], nIN = h[4];
    ok = ok && fread(coord, sizeof(double), (size_t)nN * 3, f) == (size_t)nN * 3;
    ok = ok && fread(elemToNode, sizeof(int), (size_t)nE * 4, f) == (size_t)nE * 4;
    int nl = (nI > 1 ? nI : 1) * 3;
    ok = ok && fread(neighborsList, sizeof(int), (size_t)nl, f) == (size_t)nl;
    ok = ok && fread(intfIndex, sizeof(int), (size_t)nI + 1, f) == (size_t)nI + 1;
    ok = ok && fread(intfNodes, sizeof(int), (size_t)nIN, f) == (size_t)nIN;
    ok = ok && fread(boundNodesCode, sizeof(int), (size_t)nN, f) == (size_t)nN;
    fclose(f);
    return ok ? OR_OK : OR_ERR_IO;
}

/* ------------------------------------------------------------------------- */
/* timed FEM loop: src/FEM.cc:139-257 phase structure                         */
/* ------------------------------------------------------------------------- */

static double or_now_us(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e6 + ts.tv_nsec * 1e-3;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int or_fem_loop(const double *coord, double *values, double *prec, const int *row,
                const int *col, const int *elemToNode, const int *elemToEdge,
                const double *coef, const int *checkBounds, const int *colorToElem,
                int nbColors, int nbElem, int nbNodes, int nbEdges, int operatorID,
                int nbIter, int mode, double *times_us, uint64_t *cycles)
{
    int dim = operatorID == 0 ? 1 : 9;
    double t[4] = {0, 0, 0, 0};
    uint64_t cy[4] = {0, 0, 0, 0};
    int timed = 0, nsing = 0;
    for (int it = 0; it < nbIter; it++) {
        int time_it = (nbIter == 1 || it > 0);
        double t0 = or_now_us();
        uint64_t c0 = __rdtsc();
        if (mode == 0)
            or_assembly(coord, values, row, col, elemToNode, elemToEdge, coef, nbElem,
                        nbEdges, dim, operatorID);
        else
            or_assembly_coloring(coord, values, row, col, elemToNode, elemToEdge, coef,
                                 colorToElem, nbColors, nbEdges, dim, operatorID, mode == 2);
        double t1 = or_now_us();
        uint64_t c1 = __rdtsc();
        or_prec_init(prec, values, row, col, nbNodes, dim, mode != 0);
        double t2 = or_now_us();
        uint64_t c2 = __rdtsc();
        /* halo exchange: no-op at one rank (src/halo.cc:44) */
        double t3 = or_now_us();
        uint64_t c3 = __rdtsc();
        nsing = or_prec_inversion(prec, row, col, checkBounds, nbNodes, operatorID, mode != 0);
        double t4 = or_now_us();
        uint64_t c4 = __rdtsc();
        if (time_it) {
            t[0] += t1 - t0; t[1] += t2 - t1; t[2] += t3 - t2; t[3] += t4 - t3;
            cy[0] += c1 - c0; cy[1] += c2 - c1; cy[2] += c3 - c2; cy[3] += c4 - c3;
            timed++;
        }
    }
    for (int p = 0; p < 4; p++) {
        if (times_us) times_us[p] = timed ? t[p] / timed : 0;
        if (cycles) cycles[p] = timed ? cy[p] / (uint64_t)timed : 0;
    }
    return nsing;
}
