/*
 * minifem — MI355X-native Mini-FEM assembly + block-Jacobi preconditioner.
 *
 * C ABI of libminifem.so. Plain pointers and sizes only; every function returns an
 * int status (MF_OK = 0, negative on error) unless it returns a value. Index
 * conventions are the reference's: elemToNode and CSR column ids are 1-based, CSR row
 * offsets and CSR positions are 0-based, ela blocks are 9 doubles row-major, and
 * checkBounds is markf(N,3) in Fortran order (checkBounds[k*N + i]).
 *
 * Host setup functions run on the CPU (they are the reference's one-time setup).
 * Per-iteration functions run on the GPU: pointer arguments documented "device" must
 * be device memory (hipMalloc or mf_dev_alloc); `stream` is a hipStream_t
 * passed as void* (NULL = the default stream). Device entry points return after
 * enqueueing work unless stated otherwise.
 */
#ifndef MINIFEM_H
#define MINIFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_OK 0
#define MF_ERR_ARG (-1)      /* bad argument (null pointer, negative size, bad op id) */
#define MF_ERR_IO (-2)       /* file missing, truncated or unwritable */
#define MF_ERR_COLORS (-3)   /* greedy coloring needs more than MF_MAX_COLOR colors */
#define MF_ERR_NNZ (-4)      /* CSR pattern larger than the nnz capacity given */
#define MF_ERR_HIP (-5)      /* HIP runtime error or no GPU */
#define MF_ERR_NOMEM (-6)    /* host or device allocation failed */
#define MF_ERR_BUDGET (-7)   /* request rejected by the HBM budget */
#define MF_ERR_STATE (-8)    /* object used in the wrong state */

#define MF_MAX_COLOR 128     /* src/coloring.cc:22 */
#define MF_OP_LAP 0          /* operatorID, src/main.cc:131-138 */
#define MF_OP_ELA 1
#define MF_PREC_FP64 0       /* plain FP64 (the drop-in) */
#define MF_PREC_MIXED 1      /* FP32 element products, FP64 Jacobian/scatter/inverse */
#define MF_PREC_MIXED_F32COEF 2 /* MF_PREC_MIXED taking the coefficient field as float
                                   (k[nE] lap, (lambda, mu)[nE] ela): the precision the
                                   element products use; mf_plan_assemble only */

const char *mf_status_string(int status);
int mf_version(void);

/* ---------------------------------------------------------------------------------
 * Mesh and reference-norm files (the reference's native-endian binary format).
 * ------------------------------------------------------------------------------- */

/* Header of src/IO.cc:75-80: {nbElem, nbNodes, nbEdges, nbIntf, nbIntfNodes,
 * nbBoundNodes}. Replaces the first half of read_input_data (src/IO.cc:61-80). */
int mf_read_header(const char *path, int header[6]);

/* Body of read_input_data (src/IO.cc:82-95) into caller-allocated arrays sized from
 * the header: coord[3N], elemToNode[4nE], neighborsList[3*max(nbIntf,1)],
 * intfIndex[nbIntf+1], intfNodes[nbIntfNodes], boundNodesCode[N]. */
int mf_read_input_data(const char *path, double *coord, int *elemToNode,
                       int *neighborsList, int *intfIndex, int *intfNodes,
                       int *boundNodesCode);

/* store_input_data_ (src/IO.cc:99-127), one rank: writes the header given plus the
 * arrays, to `path` (the reference writes "inputs_<np>_<rank>" in the cwd). */
int mf_store_input_data(const char *path, const int header[6], const double *coord,
                        const int *elemToNode, const int *neighborsList,
                        const int *intfIndex, const int *intfNodes,
                        const int *boundNodesCode);

/* read_ref_assembly (src/IO.cc:26-39): two numbers, values norm then prec norm. */
int mf_read_ref_assembly(const char *path, double *matrixNorm, double *precNorm);

/* store_ref_assembly_ (src/IO.cc:42-58): writes both norms with 17 digits. */
int mf_store_ref_assembly(const char *path, double matrixNorm, double precNorm);

/* ---------------------------------------------------------------------------------
 * Synthetic jittered-cube meshes (no reference counterpart: the reference reads
 * external meshes). n nodes per side, Kuhn 6-tet split of each hex, hexes and nodes
 * i-fastest, interior nodes jittered uniformly by +-jitter*h, optional seeded node
 * shuffle. Boundary codes: x=0 -> 52, y=0 -> 53, z=0 -> 54, x/y/z=1 -> 10, with
 * precedence 10 > 52 > 53 > 54.
 * ------------------------------------------------------------------------------- */

/* nbElem = 6(n-1)^3, nbNodes = n^3, nnz = N + 2(3mn^2 + 3nm^2 + m^3), m = n-1. */
int mf_cube_sizes(int n, int *nbElem, int *nbNodes, int64_t *nnz);

/* shuffleSeed == 0 keeps the natural node order. */
int mf_generate_cube(int n, double jitter, uint64_t geomSeed, uint64_t shuffleSeed,
                     double *coord, int *elemToNode, int *boundNodesCode);

/* lap: coef[e] = lognormal(0, 0.5). ela: E = lognormal(0, 0.5), nu = U[0.2, 0.35],
 * coef[2e] = lambda = E nu / ((1+nu)(1-2nu)), coef[2e+1] = mu = E / (2(1+nu)). */
int mf_generate_coef(int operatorID, int nbElem, uint64_t seed, double *coef);

/* ---------------------------------------------------------------------------------
 * Host setup (bit-exact twins of the reference's setup, src/main.cc:209-351).
 * ------------------------------------------------------------------------------- */

/* DC_create_nodeToElem (external DC-lib, called at src/main.cc:247): for each node
 * the 0-based ids of its elements in ascending order. index[N+1], value[4nE]. */
int mf_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes, int *index,
                         int *value);

/* create_nodeToNode (src/matrix.cc:55-91): row[N+1], col[nbEdges]; first-seen column
 * order, diagonal included. Returns the nnz (>= 0) or MF_ERR_NNZ when the pattern
 * does not fit in nbEdges (the reference overflows silently, src/matrix.cc:81). */
int mf_create_nodeToNode(int *row, int *col, const int *n2eIndex, const int *n2eValue,
                         const int *elemToNode, int nbNodes, int nbEdges);

/* create_elemToEdge (src/matrix.cc:25-52): elemToEdge[16 nE] CSR positions. */
int mf_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                         int *elemToEdge, int nbElem);

/* coloring_creation (src/coloring.cc:84-109): greedy first-fit over elements sharing
 * a node; colorPerm[old] = new position (stable sort by color); colorToElem has
 * MF_MAX_COLOR+1 slots (the reference's global colorToElem, src/coloring.cc:103);
 * colorPart[nE] (may be NULL) receives each element's color. */
int mf_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                         int nbElem, int nbNodes, int *colorToElem, int *nbColors);

/* DC_permute_int_2d_array (src/main.cc:229): row old moves to row perm[old]. */
int mf_permute_int_2d(int *tab, const int *perm, int n, int dim);
int mf_permute_double_2d(double *tab, const int *perm, int n, int dim);

/* dqmrd4_ + e_essbcm_ (src/main.cc:340-345; src/Fortran/qdmrd4.f,
 * src/Fortran/e_cgmelissa.F): checkBounds[3N] from boundNodesCode[N]. nbBoundNodes
 * is the header's count (the capacity DQMRD4 is given). */
int mf_boundary_mask(const int *boundNodesCode, int nbNodes, int nbBoundNodes,
                     int *checkBounds);

/* compute_double_norm (src/FEM.cc:48-56), host, serial order. */
double mf_compute_double_norm(const double *tab, int64_t size);

/* Device twins of the setup above, bit-exact with them, for meshes that are already
 * resident on the GPU (all arrays device memory; each call returns after its work is
 * done).
 *   mf_dev_create_nodeToElem: index[N+1], value[4nE], each node's elements ascending.
 *   mf_dev_create_nodeToNode: row[N+1], col[nbEdges]; returns nnz or MF_ERR_NNZ.
 *   mf_dev_create_elemToEdge: elemToEdge[16 nE].
 *   mf_dev_coloring_creation (src/coloring.cc:46-109): mf_coloring_creation's outputs (colorPart may be NULL;
 *     colorToElem in device memory, nbColors a host pointer), from the elements in
 *     dependency levels (an element is colored once every lower-index element sharing a
 *     node with it is), so the colors equal the sequential scan's. elemToNode 16-byte
 *     aligned (MF_ERR_ARG otherwise), as hipMalloc returns it. */
int mf_dev_create_nodeToElem(const int *elemToNode, int nbElem, int nbNodes, int *index,
                             int *value, void *stream);
int mf_dev_create_nodeToNode(int *row, int *col, const int *n2eIndex, const int *n2eValue,
                             const int *elemToNode, int nbNodes, int nbEdges, void *stream);
int mf_dev_create_elemToEdge(const int *row, const int *col, const int *elemToNode,
                             int *elemToEdge, int nbElem, void *stream);
int mf_dev_coloring_creation(const int *elemToNode, int *colorPart, int *colorPerm,
                             int nbElem, int nbNodes, int *colorToElem, int *nbColors,
                             void *stream);
/* mf_permute_int_2d / mf_permute_double_2d on device arrays (MF_ERR_ARG, tab unchanged,
 * when a perm entry is outside [0, n)). */
int mf_dev_permute_int_2d(int *tab, const int *perm, int n, int dim, void *stream);
int mf_dev_permute_double_2d(double *tab, const int *perm, int n, int dim, void *stream);

/* ---------------------------------------------------------------------------------
 * GPU: drop-in twins of the per-iteration entry points (device pointers).
 * ------------------------------------------------------------------------------- */

/* Number of visible GPUs (>= 0), or MF_ERR_HIP. */
int mf_device_count(void);

/* Make `device` current for this thread and check it is gfx950. */
int mf_device_init(int device);

/* Device memory, streams and events through libminifem's own HIP runtime (no
 * reference counterpart: the reference is host-only). Copies are synchronous with
 * respect to the host except mf_memcpy_d2d, which is only enqueued. */
int mf_dev_alloc(void **ptr, size_t bytes);
int mf_dev_free(void *ptr);
int mf_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int mf_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int mf_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int mf_stream_create(void **stream);
int mf_stream_destroy(void *stream);
int mf_stream_sync(void *stream);
int mf_device_sync(void);
int mf_event_create(void **event);
int mf_event_destroy(void *event);
int mf_event_record(void *event, void *stream);
int mf_event_elapsed_ms(void *start, void *stop, float *ms);

/* Pinned host memory (hipHostMalloc), asynchronous copies (the host buffer must be
 * pinned for them to overlap), event polling (done = 1 when complete) and cross-stream
 * ordering: the request loop's building blocks. */
int mf_host_alloc(void **ptr, size_t bytes);
int mf_host_free(void *ptr);
int mf_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream);
int mf_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream);
int mf_event_query(void *event, int *done);
int mf_stream_wait_event(void *stream, void *event);

/* Stream capture into an executable HIP graph: mf_graph_begin(stream), enqueue work
 * (e.g. one mf_plan_assemble per mesh of a batch) on that stream, mf_graph_end(stream,
 * &graph); mf_graph_launch replays the whole batch as one submission. The captured
 * kernels read and write the buffers they were given at capture time. */
int mf_graph_begin(void *stream);
int mf_graph_end(void *stream, void **graph);
int mf_graph_launch(void *graph, void *stream);
int mf_graph_destroy(void *graph);

/* Plain streaming device-to-device copy kernel (16-byte aligned, size a multiple of
 * 16): the HBM bandwidth reference for rooflines and byte-counter calibration. */
int mf_hbm_copy(void *dst, const void *src, size_t bytes, void *stream);

/* Measurement only: dst[i] = src[idx[i]] for i < n (device arrays; every idx in range is
 * the caller's duty). Random 8-byte reads move whole DRAM sectors: the reference for the
 * byte counters' scale on scattered reads. */
int mf_gather_probe(double *dst, const double *src, const int64_t *idx, int64_t n, void *stream);

/* assembly (src/headers/assembly.h:61-74, called at src/FEM.cc:183), COLORING
 * semantics (src/assembly.cc:593-611, :660-673): zero values[nbEdges*operatorDim],
 * then one race-free launch per color over elements [colorToElem[c], colorToElem[c+1])
 * of the color-permuted elemToNode. elemToEdge (device, 16 per element) gives the
 * scatter positions; NULL selects the reference's default row search
 * (src/assembly.cc:415-421). colorToElem is a HOST array of nbColors+1 offsets (the
 * reference's global). Device: coord, values, row, col, elemToNode, elemToEdge. */
int mf_assembly(const double *coord, double *values, const int *row, const int *col,
                const int *elemToNode, const int *elemToEdge, int nbElem, int nbEdges,
                int operatorDim, int operatorID, const int *colorToElem, int nbColors,
                void *stream);

/* mf_assembly with a per-element coefficient (device; lap: coef[nE]; ela: lambda,mu
 * pairs coef[2nE]; NULL = the reference's constants) and precision MF_PREC_*. */
int mf_assembly_ex(const double *coord, const double *coef, double *values,
                   const int *row, const int *col, const int *elemToNode,
                   const int *elemToEdge, int nbElem, int nbEdges, int operatorDim,
                   int operatorID, const int *colorToElem, int nbColors, int precision,
                   void *stream);

/* prec_init (src/preconditioner.cc:52-87, called at src/FEM.cc:199). Device. */
int mf_prec_init(double *prec, const double *values, const int *row, const int *col,
                 int nbNodes, int operatorDim, void *stream);

/* prec_inversion (src/preconditioner.cc:25-49, called at src/FEM.cc:231) with the
 * per-node ela_invert_prec of src/Fortran/elasclpr.f:2-56 (Dirichlet mask + 3x3
 * DGETRF/DGETRI). singular (device int, may be NULL) is incremented once per singular
 * ela block (the reference only prints LAPACK's info). Device pointers. */
int mf_prec_inversion(double *prec, const int *row, const int *col,
                      const int *checkBounds, int nbNodes, int operatorID,
                      int *singular, void *stream);

/* Consumer of one iteration's outputs (the reference stops before its "solver & update",
 * src/main.cc:353): block-Jacobi preconditioned CG on K = P A P + (I - P), A the
 * assembled CSR operator (values, row, col as assembled), P zeroing the dofs that
 * checkBounds masks (ela per dof; lap a node with any flag; NULL: none), M^-1 = prec as
 * prec_inversion leaves it. Solves K x = P b from the initial x (device arrays of N*d
 * doubles) to ||r|| <= rtol ||P b|| or maxIter iterations; reports both. */
int mf_pcg(int operatorID, int nbNodes, const int *row, const int *col, const double *values,
           const double *prec, const int *checkBounds, const double *b, double *x, double rtol,
           int maxIter, int *iterations, double *relResidual, void *stream);

/* mf_pcg on the system renumbered by order (device, nbNodes ints: new node t is node
 * order[t]; MF_ERR_ARG unless a permutation of 0..nbNodes-1, and for a row_ptr that does
 * not start at 0 and grow, or a column outside 1..nbNodes). values, col, prec, b and x are
 * copied into that numbering (temporary device memory of about the size of values), the
 * solve runs there, and x is copied back. An order that keeps neighbouring nodes close
 * (mf_plan_row_order) turns the operator product's scattered reads into cache hits on a
 * node-shuffled mesh. Same iterates as mf_pcg up to the order of the dot products' terms.
 * No counterpart in the reference, which has no solver. */
int mf_pcg_ordered(int operatorID, int nbNodes, const int *row, const int *col,
                   const double *values, const double *prec, const int *checkBounds,
                   const int *order, const double *b, double *x, double rtol, int maxIter,
                   int *iterations, double *relResidual, void *stream);

/* compute_double_norm (src/FEM.cc:48-56) on the device: deterministic tree sum of
 * squares, sqrt on the host. x is device memory; blocks until the result is ready. */
int mf_norm(const double *x, int64_t n, double *result, void *stream);

/* mf_norm without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same kernels and summation order as mf_norm,
 * so the same result bit for bit. Its scratch is a fixed 8 KiB array in device memory (the
 * 1024 per-workgroup partial sums), one per GPU, held until the process ends; any n >= 0 is
 * accepted, since the partial sums do not grow with n. Every call on one GPU shares that
 * GPU's scratch: calls that overlap on one GPU (several streams or host threads at once)
 * race and must be serialized by the caller. Blocks until the result is ready. */
int mf_norm_noalloc(const double *x, int64_t n, double *result, void *stream);

/* ---------------------------------------------------------------------------------
 * Plans: a mesh resident on the GPU with a precomputed tiled gather schedule.
 * One plan = one mesh topology (color-permuted elemToNode, CSR pattern, boundary
 * mask). mf_plan_assemble runs the whole per-iteration step (src/FEM.cc:183-232):
 * assembly into values + the inverted block-Jacobi prec, in one fused launch, with the
 * same per-entry summation order as the COLORING build (ascending permuted element).
 * ------------------------------------------------------------------------------- */

typedef struct mf_plan mf_plan;

/* Tuning knobs of the tiled schedule (0 = default). An off-diagonal block (c,a) is
 * the exact transpose of (a,c) (same elements, same order, commutative products), so
 * when a and c are rows of one tile the plan computes the pair once and writes both;
 * a tile computes every block of its rows itself, so each tile writes whole CSR rows. */
typedef struct mf_plan_params {
    int tile_lds_bytes;   /* LDS budget of one tile: 4096..163840 (default 155000) */
    int block_threads;    /* threads per workgroup: 512 or 1024 (default 1024) */
    int max_tile_rows;    /* cap on rows per tile, <= 256 (default 254) */
    int id_chunks;        /* > 1: bisect each of that many contiguous node-id ranges on its
                             own, so the tiles that run together write nearby CSR rows, at
                             the price of more halo elements per tile (default 1) */
    int stage_precision;  /* the precision mode the tiles are sized for. MF_PREC_FP64 (0,
                             default): the plan runs every mode. MF_PREC_MIXED or
                             MF_PREC_MIXED_F32COEF: tiles sized by the float stage, which
                             holds twice the values per byte (ela 100^3: 18 % fewer tiles);
                             the plan then runs the two mixed modes only, and FP64 launches
                             of it (mf_plan_assemble, mf_plans_assemble) return MF_ERR_ARG */
} mf_plan_params;

/* Host arrays: elemToNode (color-permuted, 1-based, 4nE), row[N+1], col[nnz],
 * checkBounds[3N] (may be NULL), coord[3N] (may be NULL; used only to group nodes into
 * tiles by coordinate bisection, else a breadth-first node order is used). The plan
 * copies what it needs to the GPU and checks every index its kernel follows. Limits: at
 * most 128 elements per node (a mesh that mf_coloring_creation accepts has no more: the
 * elements at one node need pairwise different colors) and 255 CSR entries per row
 * (MF_ERR_ARG otherwise; the colored path has neither). The
 * build runs on up to 16 host threads (environment MF_PLAN_THREADS sets the count); the
 * plan does not depend on the count. */
int mf_plan_create(mf_plan **plan, int operatorID, int nbElem, int nbNodes, int nnz,
                   const int *elemToNode, const int *row, const int *col,
                   const int *checkBounds, const double *coord,
                   const mf_plan_params *params);
int mf_plan_destroy(mf_plan *plan);

/* Builds and validates the schedule on the host only (no GPU needed) and returns the
 * statistics mf_plan_stats would report. For tests and the autotuner. */
int mf_plan_host_stats(int operatorID, int nbElem, int nbNodes, int nnz,
                       const int *elemToNode, const int *row, const int *col,
                       const double *coord, const mf_plan_params *params, int64_t out[8]);

/* Plan statistics: out[0] tiles, [1] tile-element slots, [2] off-diagonal tasks
 * (a pair computed once counts once), [3] incidence entries, [4] plan bytes on the
 * device, [5] LDS bytes per workgroup (in the stage precision), [6] threads per
 * workgroup, [7] rows per tile (max). */
int mf_plan_stats(const mf_plan *plan, int64_t out[8]);

/* The plan's node order (device, N ints, 0-based): the rows of its tiles in tile order
 * (the rows of a tile are spatially close), then the nodes with an empty row; a
 * permutation of 0..N-1, for mf_pcg_ordered. No counterpart in the reference. */
int mf_plan_row_order(const mf_plan *plan, int *order, void *stream);

/* One iteration: values (assembly, src/FEM.cc:183) and the inverted block-Jacobi prec
 * (prec_init + prec_inversion, src/FEM.cc:199-232) in one launch. Device: coord[3N],
 * coef (NULL = reference constants; points to float with MF_PREC_MIXED_F32COEF),
 * values[nnz*d], prec[N*d], singular (may be NULL; incremented per singular ela
 * block). The plan is only read: one plan may be launched from several host threads and
 * on several streams at once (each launch with its own outputs), and under graph
 * capture (no runtime queries). */
int mf_plan_assemble(mf_plan *plan, const double *coord, const double *coef,
                     double *values, double *prec, int precision, int *singular,
                     void *stream);

/* ---------------------------------------------------------------------------------
 * Several plans in one launch (no counterpart in the reference, which assembles one
 * mesh: this is the single launch of the batched and served forms).
 * ------------------------------------------------------------------------------- */

#define MF_MAX_PLAN_JOBS 16

/* One plan's iteration in a multi-plan launch: what mf_plan_assemble takes for it. */
typedef struct mf_plan_job {
    const mf_plan *plan;
    const double *coord;   /* device, 3N */
    const double *coef;    /* device, NULL = reference constants (float with F32COEF) */
    double *values;        /* device, nnz*d */
    double *prec;          /* device, N*d */
    int *singular;         /* device, may be NULL */
} mf_plan_job;

/* The iterations of 1..MF_MAX_PLAN_JOBS plans, lap and ela mixed, in ONE persistent
 * launch: the tiles of every job, one job after another, dealt to the workgroups as one
 * plan's tiles are. Each job's values and prec are bit-identical to its own
 * mf_plan_assemble. The plans must share block_threads, and their largest tiles must fit
 * one CU's LDS together with a front part sized for both operators (MF_ERR_ARG otherwise:
 * launch them in smaller groups; mf_plans_lds_bytes tells in advance). The jobs' outputs
 * must not overlap. A plan with nodes outside every element adds one small launch for
 * their prec. The job table travels in the kernel arguments, so the call can be captured
 * into a graph and the caller's table can be reused at once. */
int mf_plans_assemble(const mf_plan_job *jobs, int njobs, int precision, void *stream);

/* LDS bytes per workgroup that mf_plans_assemble would use for these jobs (0 when they
 * have no tiles), or a negative status (MF_ERR_ARG also when they do not fit one CU). */
int mf_plans_lds_bytes(const mf_plan_job *jobs, int njobs, int precision);

/* ---------------------------------------------------------------------------------
 * Served form (no counterpart in the reference): a native request loop over resident
 * meshes. Requests arrive over time (each a fresh coefficient field or fresh coordinates
 * for one mesh); each is admitted once its input and outputs fit the HBM arena the caller
 * hands over, and a free stream serves up to max_group admitted requests of one
 * precision mode that arrived together with one mf_plans_assemble launch. A request that
 * does not fit waits; later ones may pass it (first fit) until it has been passed
 * max_bypass times.
 * ------------------------------------------------------------------------------- */

typedef struct mf_server mf_server;

#define MF_REQ_COEF 0   /* the payload is the mesh's coefficient field */
#define MF_REQ_COORD 1  /* the payload is the mesh's coordinates */

/* A resident mesh: its plan and device inputs (used where a request does not replace
 * them), and its sizes: a request's payload must be exactly ncoord doubles
 * (coordinates), ncoef doubles (a coefficient field) or ncoef floats (a coefficient field
 * with MF_PREC_MIXED_F32COEF). */
typedef struct mf_server_mesh {
    const mf_plan *plan;
    const double *coord;   /* device, 3N */
    const double *coef;    /* device, NULL = reference constants */
    int64_t nvalues;       /* nnz * d */
    int64_t nprec;         /* N * d */
    int64_t ncoord;        /* 3N */
    int64_t ncoef;         /* nE (lap) or 2 nE (ela) */
} mf_server_mesh;

typedef struct mf_request {
    int mesh;              /* index of a resident mesh */
    int kind;              /* MF_REQ_COEF or MF_REQ_COORD */
    int mode;              /* precision of its launch, MF_PREC_*; MF_PREC_MIXED_F32COEF only
                              for a float coefficient field */
    int pad;
    const void *payload;   /* host memory (pinned, so its upload overlaps other work) */
    int64_t payload_bytes;
    double t_arrival;      /* seconds after the start of the run */
    double *values_out;    /* host, nvalues doubles, or NULL: copied back once served */
    double *prec_out;      /* host, nprec doubles, or NULL */
} mf_request;

typedef struct mf_server_stats {
    int64_t launches;
    int64_t memory_waits;  /* requests that found a stream free but not their HBM */
    double makespan_s;     /* first arrival to last completion */
} mf_server_stats;

/* arena: device memory of arena_bytes that the server carves for request inputs and
 * outputs (no allocation while serving); slots: streams (1..16); max_group:
 * 1..MF_MAX_PLAN_JOBS; max_bypass >= 0. */
int mf_server_create(mf_server **server, void *arena, int64_t arena_bytes, int slots,
                     int max_group, int max_bypass);
/* Adds a resident mesh, also between runs; returns its index (>= 0) or a status. */
int mf_server_add_mesh(mf_server *server, const mf_server_mesh *mesh);
/* Serves n requests (any order of t_arrival) and returns when all are done: t_start[i]
 * (admission) and t_done[i] (the end of its launch and copies, as the loop saw it), in
 * seconds after the start. MF_ERR_BUDGET, before anything runs, when a request needs
 * more than the arena; MF_ERR_ARG for a bad mesh, kind, mode or payload size, or a mode
 * the mesh's plan does not launch in (an FP64 request on a plan staged in float). */
int mf_server_run(mf_server *server, const mf_request *reqs, int n, double *t_start,
                  double *t_done, mf_server_stats *stats);
/* One run at a time per server (the loop is single-threaded; two threads running one
 * server race). Arena bytes held by requests (0 between runs): */
int64_t mf_server_arena_used(const mf_server *server);
int mf_server_destroy(mf_server *server);

/* Measurement only, no counterpart in the reference: writes zeros over values[nnz*d]
 * with the assembly's output stream alone, to time what HBM accepts for it. mode 0:
 * every tile row's CSR segment in the order mf_plan_assemble writes them; mode 1: the
 * whole array front to back (the same bytes); mode 2: only the 128-byte lines each
 * segment covers whole; mode 3: only the partial lines at the segments' ends. The values
 * must be assembled again afterwards. */
int mf_plan_write_probe(const mf_plan *plan, double *values, int mode, void *stream);

#ifdef __cplusplus
}
#endif
#endif
