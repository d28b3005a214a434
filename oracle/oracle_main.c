/*
 * Host-core baseline binary built from the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Mirrors the reference driver's setup pipeline and FEM loop at one rank
 * (src/main.cc:131-351, src/FEM.cc:139-261) so that the reference's `ref` and
 * OpenMP `coloring` builds can be timed on the host cores:
 *
 *   minifem_oracle <mesh-file> <lap|ela> <iters> <ref|coloring|coloring-par>
 *                  [coef-file]
 *
 * The mesh file is the reference's binary input format (src/IO.cc:61-96). The
 * optional coefficient file holds raw float64: nbElem (lap) or 2*nbElem (ela, lambda
 * then mu per element) in the file's element order. Prints per-phase microseconds and
 * RDTSC cycles (iteration 0 excluded when iters > 1, src/FEM.cc:182) and the norms
 * of values and prec (src/FEM.cc:59-98).
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <mesh-file> <lap|ela> <iters> <ref|coloring|coloring-par> [coef-file]\n", argv[0]);
        return 1;
    }
    const char *path = argv[1];
    int opID = strcmp(argv[2], "ela") == 0 ? 1 : 0;
    if (strcmp(argv[2], "lap") && strcmp(argv[2], "ela")) { fprintf(stderr, "bad operator\n"); return 1; }
    int iters = atoi(argv[3]);
    if (iters < 1) { fprintf(stderr, "Number of iterations must be at least 1.\n"); return 1; }
    int mode = !strcmp(argv[4], "ref") ? 0 : !strcmp(argv[4], "coloring") ? 1 : 2;
    int dim = opID ? 9 : 1;

    int h[6];
    if (or_read_header(path, h)) { fprintf(stderr, "cannot read %s\n", path); return 1; }
    int nE = h[0], nN = h[1], nnzCap = h[2], nI = h[3], nIN = h[4];
    double *coord = malloc(sizeof(double) * (size_t)nN * 3);
    int *e2n = malloc(sizeof(int) * (size_t)nE * 4);
    int *nl = malloc(sizeof(int) * (size_t)((nI > 1 ? nI : 1) * 3));
    int *ii = malloc(sizeof(int) * (size_t)(nI + 1));
    int *in = malloc(sizeof(int) * (size_t)(nIN + 1));
    int *code = malloc(sizeof(int) * (size_t)nN);
    if (or_read_input_data(path, coord, e2n, nl, ii, in, code)) { fprintf(stderr, "read error\n"); return 1; }
    double *coef = NULL;
    if (argc > 5) {
        size_t nc = (size_t)nE * (opID ? 2 : 1);
        coef = malloc(sizeof(double) * nc);
        FILE *f = fopen(argv[5], "rb");
        if (!f || fread(coef, sizeof(double), nc, f) != nc) { fprintf(stderr, "coef read error\n"); return 1; }
        fclose(f);
    }

    double t0 = now_s();
    int colorToElem[OR_MAX_COLOR + 1], nColors = 1;
    colorToElem[0] = 0; colorToElem[1] = nE;
    if (mode != 0) {
        int *perm = malloc(sizeof(int) * (size_t)nE);
        if (or_coloring_creation(e2n, NULL, perm, nE, nN, colorToElem, &nColors)) {
            fprintf(stderr, "Error: Not enough colors.\n"); return 1;
        }
        or_permute_int_2d(e2n, perm, nE, 4);
        if (coef) or_permute_double_2d(coef, perm, nE, opID ? 2 : 1);
        free(perm);
    }
    double t1 = now_s();
    int *n2eI = malloc(sizeof(int) * (size_t)(nN + 1));
    int *n2eV = malloc(sizeof(int) * (size_t)nE * 4);
    int *row = malloc(sizeof(int) * (size_t)(nN + 1));
    int *col = malloc(sizeof(int) * (size_t)nnzCap);
    or_create_nodeToElem(e2n, nE, nN, n2eI, n2eV);
    int nnz = or_create_nodeToNode(row, col, n2eI, n2eV, e2n, nN, nnzCap);
    if (nnz < 0) { fprintf(stderr, "CSR larger than the header nnz\n"); return 1; }
    free(n2eI); free(n2eV);
    double t2 = now_s();
    int nb = h[5];
    int *blist = malloc(sizeof(int) * (size_t)(nb + 1));
    int *cb = malloc(sizeof(int) * (size_t)nN * 3);
    or_dqmrd4(nN, code, &nb, blist);
    or_e_essbcm(nN, nb, blist, code, cb);
    double t3 = now_s();

    double *values = malloc(sizeof(double) * (size_t)nnz * dim);
    double *prec = malloc(sizeof(double) * (size_t)nN * dim);
    double tus[4];
    uint64_t cyc[4];
    double t4 = now_s();
    int nsing = or_fem_loop(coord, values, prec, row, col, e2n, NULL, coef, cb, colorToElem,
                            nColors, nE, nN, nnz, opID, iters, mode, tus, cyc);
    double t5 = now_s();
    printf("{\"mode\": \"%s\", \"op\": \"%s\", \"threads\": %d, \"nbElem\": %d, \"nbNodes\": %d, "
           "\"nnz\": %d, \"colors\": %d, \"setup_s\": {\"coloring\": %.3f, \"csr\": %.3f, \"bounds\": %.3f}, "
           "\"loop_s\": %.3f, \"phase_us\": [%.1f, %.1f, %.1f, %.1f], "
           "\"phase_cycles\": [%llu, %llu, %llu, %llu], \"singular\": %d, "
           "\"values_norm\": %.17g, \"prec_norm\": %.17g}\n",
           argv[4], argv[2], mode ? or_num_threads() : 1, nE, nN, nnz, nColors, t1 - t0, t2 - t1,
           t3 - t2, t5 - t4, tus[0], tus[1], tus[2], tus[3], (unsigned long long)cyc[0],
           (unsigned long long)cyc[1], (unsigned long long)cyc[2], (unsigned long long)cyc[3],
           nsing, or_compute_double_norm(values, (int64_t)nnz * dim),
           or_compute_double_norm(prec, (int64_t)nN * dim));
    return 0;
}
