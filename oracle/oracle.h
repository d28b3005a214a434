/*
 * CPU oracle for Mini-FEM's per-iteration assembly and block-Jacobi preconditioner.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference algorithm,
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker and as the timed host-core baseline. Nothing in the product library links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * Mini-FEM source tree). Conventions follow the reference exactly:
 *   - elemToNode and CSR column ids are 1-based, CSR row offsets are 0-based
 *     (src/assembly.cc:91, src/matrix.cc:65, src/matrix.cc:81);
 *   - CSR columns are in first-seen order, diagonal included (src/matrix.cc:55-91);
 *   - the ela block of a CSR entry is 9 doubles, row-major (src/assembly.cc:386-409);
 *   - checkBounds is markf(N,3) in Fortran order: checkBounds[k*N + i]
 *     (src/Fortran/e_cgmelissa.F:1-51).
 *
 * Pinning status of each function (what independent check holds it) is recorded in
 * the header comment of oracle.c next to each function.
 */
#ifndef MINIFEM_ORACLE_H
#define MINIFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_DIM_ELEM 4
#define OR_DIM_NODE 3
#define OR_VALUES_PER_ELEM 16
#define OR_MAX_COLOR 128

/* status codes (the reference calls exit(); the oracle returns instead) */
#define OR_OK 0
#define OR_ERR_NNZ (-4)      /* CSR pattern larger than the nnz given in the header */
#define OR_ERR_COLORS (-3)   /* greedy coloring needs more than 128 colors */
#define OR_ERR_IO (-2)
#define OR_ERR_ARG (-1)

/* --- setup --------------------------------------------------------------- */
int or_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes,
                         int *n2eIndex, int *n2eValue);
int or_create_nodeToNode(int *row, int *col, const int *n2eIndex, const int *n2eValue,
                         const int *elemToNode, int nbNodes, int nbEdgesCap);
void or_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                          int *elemToEdge, int nbElem);
int or_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                         int nbElem, int nbNodes, int *colorToElem, int *nbColors);
void or_permute_int_2d(int *tab, const int *perm, int n, int dim);
void or_permute_double_2d(double *tab, const int *perm, int n, int dim);
int or_dqmrd4(int nbNodes, const int *code, int *nbBound, int *boundList);
void or_e_essbcm(int nbNodes, int nbBound, const int *boundList, const int *code,
                 int *checkBounds);

/* --- per-iteration hot path ---------------------------------------------- */
void or_elem_coef(double g[4][3], const double *coord, const int *elemToNode, int elem);
void or_assembly_lap_seq(const double *coord, double *values, const int *row,
                         const int *col, const int *elemToNode, const int *elemToEdge,
                         const double *coef, int firstElem, int lastElem);
void or_assembly_ela_seq(const double *coord, double *values, const int *row,
                         const int *col, const int *elemToNode, const int *elemToEdge,
                         const double *coef, int firstElem, int lastElem);
void or_assembly(const double *coord, double *values, const int *row, const int *col,
                 const int *elemToNode, const int *elemToEdge, const double *coef,
                 int nbElem, int nbEdges, int operatorDim, int operatorID);
void or_assembly_coloring(const double *coord, double *values, const int *row,
                          const int *col, const int *elemToNode, const int *elemToEdge,
                          const double *coef, const int *colorToElem, int nbColors,
                          int nbEdges, int operatorDim, int operatorID,
                          int parallelElements);
void or_prec_init(double *prec, const double *values, const int *row, const int *col,
                  int nbNodes, int operatorDim, int parallel);
int or_ela_invert_prec(int nbNodes, const int *row, const int *col, double *prec,
                       const int *checkBounds, int curNode1);
int or_prec_inversion(double *prec, const int *row, const int *col,
                      const int *checkBounds, int nbNodes, int operatorID, int parallel);
double or_compute_double_norm(const double *tab, int64_t size);

/* --- mixed precision restatement (FP32 element products, FP64 Jacobian) --- */
void or_assembly_mixed(const double *coord, double *values, const int *row,
                       const int *col, const int *elemToNode, const int *elemToEdge,
                       const double *coef, int nbElem, int nbEdges, int operatorDim,
                       int operatorID);

/* --- mesh file ------------------------------------------------------------ */
int or_read_header(const char *path, int hdr[6]);
int or_read_input_data(const char *path, double *coord, int *elemToNode,
                       int *neighborsList, int *intfIndex, int *intfNodes,
                       int *boundNodesCode);

/* --- timed FEM loop (CPU baseline) --------------------------------------- */
/* mode: 0 = ref (serial), 1 = coloring-omp as written (serial element loop,
 * parallel zero and prec loops), 2 = coloring-omp with the element loop parallel
 * inside each color. times_us[4] receives the average per-phase microseconds over
 * the timed iterations (iteration 0 excluded when nbIter > 1). */
int or_fem_loop(const double *coord, double *values, double *prec, const int *row,
                const int *col, const int *elemToNode, const int *elemToEdge,
                const double *coef, const int *checkBounds, const int *colorToElem,
                int nbColors, int nbElem, int nbNodes, int nbEdges, int operatorID,
                int nbIter, int mode, double *times_us, uint64_t *cycles);
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
