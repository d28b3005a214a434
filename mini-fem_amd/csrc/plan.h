// Tiled gather schedule for one mesh topology (host-built, device-resident).
//
// Rows (nodes) are grouped into tiles by recursive coordinate bisection. A tile's
// workgroup
//   A)  loads the coordinates of every node its elements touch into LDS, then computes
//       the scaled gradients (and coefficients) of every element incident to one of its
//       rows, once per tile, into LDS;
//   B1) "diagonal chunks": one lane per tile row sums the row's diagonal block over all
//       its incident elements, in ascending (color-permuted) element order -- the
//       summation order of the reference's COLORING build -- keeps it in LDS and inverts
//       it into the block-Jacobi prec (prec_init + prec_inversion fused);
//   B2) "row chunks": whole rows packed into 64-lane waves, one lane per CSR entry in
//       CSR order; an off-diagonal lane sums the elements containing both its row and
//       its column node (same order), the diagonal lane copies its block from LDS, and
//       the wave writes complete row slices, consecutive lanes on consecutive blocks.
// Every CSR entry is written exactly once; values need no zero-fill.
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

constexpr int kTaskLanes = 64;           // tasks per chunk = lanes per wave
constexpr int kDiagGroup = 4;            // diagonal-chunk contributions per 8-byte load
constexpr int kRowGroup = 2;             // row-chunk contributions per 4-byte load
constexpr int kRowPrefetch = 4;          // row-chunk groups loaded before the stores
constexpr uint32_t kNoTask = 0xFFFFFFFFu;
constexpr uint32_t kDiagLane = 1u << 16; // row-chunk task flag: copy the diagonal block
constexpr int kMaxTileElems = 4094;      // contribution entries carry a 12-bit element index
constexpr int kMaxTileNodes = 256;       // tile-local vertex ids are 8-bit
constexpr int kMaxTileRows = 256;        // task words carry 8-bit rows
constexpr int kMaxRowLen = 256;          // and 8-bit offsets inside a row
constexpr int kMaxLdsBytes = 160 * 1024; // LDS of one gfx950 CU

// Element-stage slot of one tile element, in units of the stage type (double or
// float): the (x,y) gradient pairs of vertices 0..3 at 2i (16-byte aligned for
// double2 reads), the z components at 8+i, then the coefficients at 12.. (ela:
// lambda+2mu, lambda+mu, mu; lap: k), padded to an even count.
constexpr int kStageEla = 18;
constexpr int kStageLap = 14;

// LDS of one tile, in this order (offsets 16-byte aligned where a 16-byte access needs
// it): [element stage: n_elem+1 slots, the last one all zero and the target of padding
// contributions] [diagonal blocks: n_rows] [prec blocks: n_rows] [node coordinates:
// n_node*24] [row offsets, row lengths, Dirichlet dof masks: n_rows*4 each].
struct TileLds {
    int stage, diag, prec, coords, rowptr, rowlen, mask, total;
};

MF_HOST_DEVICE inline TileLds tile_lds_layout(int n_node, int n_elem, int n_rows,
                                              int slot_bytes, int block_bytes)
{
    TileLds L;
    L.stage = 0;
    L.diag = ((n_elem + 1) * slot_bytes + 15) & ~15;
    L.prec = L.diag + n_rows * block_bytes;
    L.coords = (L.prec + n_rows * block_bytes + 15) & ~15;
    L.rowptr = L.coords + n_node * 24;
    L.rowlen = L.rowptr + n_rows * 4;
    L.mask = L.rowlen + n_rows * 4;
    L.total = L.mask + n_rows * 4;
    return L;
}

struct TileHdr {
    int32_t elem_start, n_elem;      // into tile_elems / tile_verts
    int32_t node_start, n_node;      // into tile_nodes; the first n_rows are the tile's rows
    int32_t n_rows;
    int32_t dchunk_start, n_dchunks; // diagonal chunks (into chunks; tasks at [c*64 ...])
    int32_t rchunk_start, n_rchunks; // row chunks
    int32_t rinfo_start;             // into tile_rowinfo: n_rows entries
    int32_t pad[2];
};

// Static per-tile-row data: the row's CSR offset and (length | Dirichlet mask << 16),
// mask bit k set when dof k of the node is constrained (ela only).
struct RowInfo {
    int32_t rowptr;
    int32_t len_mask;
};

// Contributions of a chunk are stored in groups of G steps (G = kDiagGroup for diagonal
// chunks, kRowGroup for row chunks): group g is 64 lanes x G entries, lane l's entries
// contiguous, so one load per lane fetches its next G contributions and a wave's load is
// one contiguous block. Step s of lane l is at contrib_start + (s/G)*64G + l*G + s%G.
// An entry is t<<4 | a<<2 | k (tile element t, its local vertices of the row and of the
// column node); padding entries are n_elem<<4, the all-zero slot.
// Diagonal-chunk task word: the tile row. Row-chunk task word: ra | oa<<8 (| kDiagLane
// for the row's diagonal entry), oa the entry's offset inside row ra.
struct ChunkHdr {
    int32_t contrib_start;
    int32_t steps;
};

struct PlanHost {
    int op = 0, nE = 0, N = 0, nnz = 0;
    std::vector<TileHdr> tiles;
    std::vector<int32_t> tile_elems;   // global (permuted) element ids, ascending per tile
    std::vector<uint8_t> tile_verts;   // 4 tile-local node ids per tile element
    std::vector<int32_t> tile_nodes;   // global node ids (0-based)
    std::vector<RowInfo> tile_rowinfo; // per tile row
    std::vector<ChunkHdr> chunks;
    std::vector<uint32_t> tasks;       // 64 task words per chunk
    std::vector<uint16_t> contribs;
    std::vector<int32_t> isolated;     // nodes with an empty row (no element)
    int max_node = 0, max_elem = 0, max_rows = 0;
    int64_t total_tasks = 0;
};

// Builds the schedule. coord (3N doubles) may be null: nodes are then ordered by a
// breadth-first sweep of the node graph instead of coordinate bisection.
// slot_bytes: LDS bytes of one element-stage slot of the widest variant that will run
// it; tiles are grown while their LDS layout stays <= lds_budget.
int build_plan(PlanHost &p, int op, int nE, int N, int nnz, const int *e2n, const int *row,
               const int *col, const int *checkBounds, const double *coord, int lds_budget,
               int max_rows, int slot_bytes);

// Checks every index the kernel will follow against the array it indexes and that every
// CSR entry is covered exactly once. Returns 0 or a negative status.
int validate_plan(const PlanHost &p, const int *row, const int *col, int slot_bytes);

}  // namespace mf
