// Tiled gather schedule for one mesh topology (host-built, device-resident).
//
// The nodes that own a CSR row are grouped into tiles by recursive coordinate bisection.
// A tile's workgroup has two kinds of waves. Loader waves issue every global load: the
// next tile's static data and its coordinate and coefficient gathers, held in registers
// and written into LDS between phases; they never store. Compute waves never load from
// global memory; they
//   A) compute the scaled gradients and coefficients of every element incident to one of
//      the tile's rows, once per tile, into an LDS element stage;
//   B) sum every CSR block of the tile's rows from the stage, each block in ascending
//      (color-permuted) element order -- the summation order of the reference's
//      COLORING build, src/assembly.cc:246-271:
//      - a row's incident elements are kept in LDS as an ascending "incidence list";
//      - an off-diagonal block (row a, column c) is one lane, which reads the elements
//        holding both a and c, with a's and c's vertex slots in them, from its task
//        record (or, past 6 elements, walks a mask over row a's incidence list);
//        block (c,a) is its exact transpose (same elements in the same order, and
//        IEEE products commute), so a pair is computed once and written twice when
//        c is a row of the same tile (a column of another tile: that tile computes
//        (c,a) itself);
//      - a diagonal block (the whole incidence list) is one lane;
//   O) put every block of the tile's rows into an LDS output image of those rows' CSR
//      segments, laid over the element stage;
//   W) copy the image row by row to the CSR value array in consecutive 16-byte pieces,
//      while the diagonal lanes invert their blocks into the block-Jacobi prec
//      (prec_init + prec_inversion fused). Meanwhile the loader waves write the next
//      tile's coordinates, coefficients and element vertices into LDS.
// Every CSR entry and prec block is written exactly once; values need no zero-fill.
// Keeping loads and stores in different waves matters on gfx950: a wave's vmcnt retires
// its loads and stores in issue order, so a wave that waited for a load issued after its
// own stores would wait for those stores too.
#pragma once

#include <cstdint>
#include <vector>

// The tile LDS layout is computed by the host (sizing) and by the kernel (offsets).
#if defined(__HIPCC__)
#define MF_HOST_DEVICE __host__ __device__
#else
#define MF_HOST_DEVICE
#endif

namespace mf {

constexpr int kLanes = 64;               // lanes per wave = CSR blocks per work block
constexpr int kMaxTileElems = 1024;      // staged elements per tile (the zero slot is extra)
constexpr int kMaxTileNodes = 256;       // tile-local node ids are 8-bit
constexpr int kMaxTileRows = 256;        // tile-local row ids are 8-bit
constexpr int kMaxTileInc = 4096;        // incidence entries per tile
// Incident elements per row. The elements holding one node share it, so greedy coloring
// gives them pairwise different colors: a mesh that colors within 128 colors
// (src/coloring.cc:22) has at most 128 per node.
constexpr int kMaxValence = 128;
constexpr int kMaskValence = 64;         // rows up to this valence: mask records (64 bits)
constexpr int kMaxLdsBytes = 160 * 1024; // LDS of one gfx950 CU
constexpr int kIncAlign = 4;             // row incidence lists start at multiples of 4
constexpr int kXcds = 8;                 // MI355X: workgroups are dealt to 8 XCDs round-robin
constexpr int kListLen = 6;              // contributions held in one task record
constexpr uint32_t kMaskRecord = 0xFFFFFFFFu;  // first word of a mask task record
constexpr uint32_t kWalkRecord = 0xFFFFFFFEu;  // first word of a walk task record

// Waves of a workgroup of `block` threads (512 or 1024): loaders, then compute waves.
MF_HOST_DEVICE constexpr int loader_waves(int block) { return block >= 1024 ? 4 : 2; }
MF_HOST_DEVICE constexpr int compute_waves(int block) { return block / kLanes - loader_waves(block); }
// Staged elements per tile: at most two per compute thread.
MF_HOST_DEVICE constexpr int max_tile_elems(int block)
{
    return kMaxTileElems < 2 * kLanes * compute_waves(block) ? kMaxTileElems
                                                               : 2 * kLanes * compute_waves(block);
}
// Work blocks of 64 lanes per tile. A diagonal block takes a compute wave; the other
// compute waves (at least one, which also copies the output out while the diagonal waves
// invert) take up to task_blocks_per_wave task blocks each, their results held in
// registers until the output image is written: two, for lap (one double per lane) and for
// ela (a 3x3 block per lane; two of them fit the 128 registers of a 1024-thread
// workgroup's lane, and larger ela tiles then stage fewer halo elements).
MF_HOST_DEVICE constexpr int task_blocks_per_wave(bool) { return 2; }
MF_HOST_DEVICE constexpr bool tile_blocks_fit(bool ela, int block, int n_diag, int n_task_blocks)
{
    return n_diag < compute_waves(block) &&
           n_task_blocks <= task_blocks_per_wave(ela) * (compute_waves(block) - n_diag);
}
// Task records one tile holds at most.
MF_HOST_DEVICE constexpr int max_tile_tasks(bool ela, int block)
{
    return task_blocks_per_wave(ela) * compute_waves(block) * kLanes;
}

// Element-stage slot of one tile element, in units of the stage type (double or float).
//   ela: vertex v's {gx, gy, gz, lambda+2mu} at 4v, then {lambda+mu, mu} at 16: a vertex's
//        gradient with the first coefficient is two 16-byte LDS reads in double, one in
//        float. 144 bytes in double; 80 in float (16-byte multiples keep those reads
//        aligned).
//   lap: the (x,y) gradient pairs at 2v, the z components at 8+v, k at 12: 112 bytes in
//        double, 56 in float.
// Every length is an odd multiple of 16 (double) or 8 (float) bytes: consecutive elements
// start in different LDS bank groups, so the per-element stage writes are conflict-free
// and the gathers of scattered elements spread over all banks.
constexpr int kStageEla64 = 18, kStageEla32 = 20;
constexpr int kStageLap = 14;
MF_HOST_DEVICE constexpr int stage_len(bool ela, bool fp64)
{
    return ela ? (fp64 ? kStageEla64 : kStageEla32) : kStageLap;
}

MF_HOST_DEVICE inline int align16(int x) { return (x + 15) & ~15; }

// LDS of one tile, in two parts.
// The front part has the same offsets for every tile of a plan (sized by its largest
// tile, f_node nodes and f_elem elements): the loader waves write the next tile's data
// there while the compute waves still read the back part of the current tile.
//   [node coordinates: f_node x 24 B] [element vertices: f_elem + 1 x 4 B]
//   [raw coefficients: f_elem x 16 B (ela: lambda, mu) or 8 B (lap: k)]
// The back part follows the tile's own sizes:
//   [overlay: the element stage (n_elem + 1 slots, the last one all zero) during A and B,
//             the output image (n_ent blocks of 9 doubles, ela, or 1, lap, in tile entry
//             order) during O and W]
//   [row info: n_rows x 16 B] [task records: n_task x 16 B] [row node ids: n_rows x 4 B]
//   [incidence lists: n_inc x 2 B, n_inc a multiple of 8].
struct TileLds {
    int coords, verts, coef, stage, rinfo, tasks, rnode, inc, total;
};

// front_ela: the front's coefficient region holds ela's 16 bytes per element (a launch
// that mixes operators uses it for every tile); ela: the operator of the tile itself.
MF_HOST_DEVICE inline TileLds tile_lds_layout(bool front_ela, bool ela, int f_node, int f_elem,
                                              int n_elem, int n_rows, int n_inc, int n_task,
                                              int n_ent, int slot_bytes)
{
    TileLds L;
    L.coords = 0;
    L.verts = align16(f_node * 24);
    L.coef = align16(L.verts + (f_elem + 1) * 4);
    L.stage = L.coef + align16(f_elem * (front_ela ? 16 : 8));
    const int stage = (n_elem + 1) * slot_bytes, out = n_ent * (ela ? 72 : 8);
    L.rinfo = L.stage + align16(stage > out ? stage : out);
    L.tasks = L.rinfo + n_rows * 16;
    L.rnode = L.tasks + n_task * 16;
    L.inc = align16(L.rnode + n_rows * 4);
    L.total = L.inc + align16(n_inc * 2);
    return L;
}

// The layout of a launch of one operator (the front sized for that operator).
MF_HOST_DEVICE inline TileLds tile_lds_layout(bool ela, int f_node, int f_elem, int n_elem,
                                              int n_rows, int n_inc, int n_task, int n_ent,
                                              int slot_bytes)
{
    return tile_lds_layout(ela, ela, f_node, f_elem, n_elem, n_rows, n_inc, n_task, n_ent,
                           slot_bytes);
}

// Diagonal work blocks of a tile: one row per lane.
MF_HOST_DEVICE inline int diag_blocks(bool, int n_rows) { return (n_rows + kLanes - 1) / kLanes; }

// Work blocks of a tile on its compute waves: diagonal block i (rows 64i..) on compute
// wave i; task block m (tasks 64m..) on compute wave diag_blocks + m % w, as its
// (m / w)-th task block, w = compute waves - diag_blocks.

struct TileHdr {
    int32_t elem_start, n_elem;  // into tile_elems / tile_verts (ascending element ids)
    int32_t node_start, n_node;  // into tile_nodes; the first n_rows are the tile's rows
    int32_t n_rows;
    int32_t rinfo_start;         // into tile_rowinfo: n_rows entries
    int32_t inc_start, n_inc;    // into inc (multiples of 8)
    int32_t task_start, n_task;  // into task_rec (4 words per task)
    int32_t n_ent;               // CSR entries of the tile's rows
    int32_t pad;
};

// Static per-tile-row data (int4 in LDS):
//   rowptr:   CSR position of the row's first entry;
//   len_diag: row length | diagonal offset << 16 | Dirichlet mask << 24 (bit k set when
//             dof k of the node is constrained, ela only);
//   inc:      tile-local offset of the row's incidence list | valence << 16;
//   ent:      tile-local index of the row's first entry (rows in tile order, each row's
//             entries in CSR order).
struct RowInfo {
    int32_t rowptr;
    int32_t len_diag;
    int32_t inc;
    int32_t ent;
};

// Incidence entry: t << 2 | a, t the tile element, a the row node's vertex slot in it.
// Each row's list starts at a multiple of kIncAlign entries (zero padding between lists:
// element 0, read but never summed), so a lane reads 4 entries with one 8-byte access.
// Off-diagonal tasks of a tile, in tile-row order and CSR order inside a row, one 16-byte
// record each (task_rec, 4 words), for the block (a,c) of tile row a and column c:
//   words 0-2: the n elements holding both a and c. n <= kListLen: six u16 contributions
//              t << 4 | a << 2 | k (tile element, slots of a and c in it) ascending, the
//              unused ones t = n_elem (the zero slot). n > kListLen and row a's valence
//              <= kMaskValence: word 0 = kMaskRecord, words 1-2 a 64-bit mask, bit b set
//              when the b-th element of row a's incidence list holds c. n > kListLen on a
//              longer row: word 0 = kWalkRecord, words 1-2 zero; the lane walks row a's
//              whole incidence list and takes the elements whose vertices include c.
//   word 3:    ra | oa << 8 | cl << 16 | oc << 24: a's tile row, the block's offset in
//              a's row, c's tile-local node, and (when cl < n_rows: c is a row of the same
//              tile, and the task also writes the transpose (c,a)) the offset of (c,a) in
//              c's row. A column of another tile: that tile computes (c,a) itself.
struct PlanHost {
    int op = 0, nE = 0, N = 0, nnz = 0, block = 512;
    std::vector<TileHdr> tiles;
    std::vector<int32_t> tile_elems;   // global (permuted) element ids, ascending per tile
    std::vector<uint8_t> tile_verts;   // 4 tile-local node ids per tile element
    std::vector<int32_t> tile_nodes;   // global node ids (0-based)
    std::vector<RowInfo> tile_rowinfo; // per tile row
    std::vector<uint16_t> inc;
    std::vector<uint32_t> task_rec;    // 4 words per task
    std::vector<int32_t> isolated;     // nodes with an empty row (no element)
    int max_node = 0, max_elem = 0, max_rows = 0;
};

// Builds the schedule. coord (3N doubles) may be null: nodes are then ordered by a
// breadth-first sweep of the node graph instead of coordinate bisection. slot_bytes:
// LDS bytes of one element-stage slot of the widest variant that will run it; tiles are
// grown while their LDS layout stays <= lds_budget and the kernel's per-thread limits
// for a workgroup of `block` threads hold. checkBounds (3N, ela) may be null.
// id_chunks > 1: the rows are first cut into that many contiguous node-id ranges, and
// each range is bisected on its own (tiles then hold rows of one id range, so the tiles
// that run at the same time write nearby CSR rows).
int build_plan(PlanHost &p, int op, int nE, int N, int nnz, const int *e2n, const int *row,
               const int *col, const int *checkBounds, const double *coord, int lds_budget,
               int max_rows, int slot_bytes, int block, int id_chunks = 1);

// Checks every index the kernel will follow against the array it indexes, that every
// CSR entry is written exactly once (diagonal, task or transpose) and that every task's
// mask selects exactly the elements holding both its nodes. Returns 0 or a negative
// status.
int validate_plan(const PlanHost &p, const int *row, const int *col, const int *e2n,
                  int slot_bytes);


}  // namespace mf
