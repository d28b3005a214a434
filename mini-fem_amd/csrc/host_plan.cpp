// Host construction and validation of the tiled gather schedule (layout in plan.h).
// The partition and the per-tile emission run on up to 16 host threads (MF_PLAN_THREADS
// overrides the count); the plan does not depend on the thread count: subtrees of the
// bisection and chunks of tiles are joined back in order.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

#include "host_threads.h"
#include "minifem.h"
#include "plan.h"

namespace mf {

namespace {

int plan_threads() { return host_threads(); }

// Map from non-negative ids to ints for the ids of one tile (a few hundred nodes, up to
// ~1K elements), open addressing, cleared in O(1) by a generation counter. A map that
// would pass half full reports it (insert returns -2) instead of degrading.
class SmallMap {
public:
    explicit SmallMap(int log2cap)
        : shift_(32 - log2cap), mask_((1u << log2cap) - 1), key_(1u << log2cap),
          val_(1u << log2cap), gen_(1u << log2cap, 0)
    {
    }
    void clear()
    {
        n_ = 0;
        if (++cur_ == 0) {
            std::fill(gen_.begin(), gen_.end(), 0u);
            cur_ = 1;
        }
    }
    int size() const { return n_; }
    // 1: inserted, 0: already present (value kept), -2: full
    int insert(int k, int v)
    {
        uint32_t h = slot(k);
        while (gen_[h] == cur_) {
            if (key_[h] == k) return 0;
            h = (h + 1) & mask_;
        }
        if (2 * (n_ + 1) > (int)mask_ + 1) return -2;
        gen_[h] = cur_;
        key_[h] = k;
        val_[h] = v;
        n_++;
        return 1;
    }
    // the value of k, or -1 when absent
    int get(int k) const
    {
        uint32_t h = slot(k);
        while (gen_[h] == cur_) {
            if (key_[h] == k) return val_[h];
            h = (h + 1) & mask_;
        }
        return -1;
    }

private:
    uint32_t slot(int k) const { return ((uint32_t)k * 2654435761u) >> shift_ & mask_; }
    int shift_;
    uint32_t mask_;
    std::vector<int> key_, val_;
    std::vector<uint32_t> gen_;
    uint32_t cur_ = 1;
    int n_ = 0;
};

// Per-thread scratch of the partitioner's fit test.
struct FitScratch {
    SmallMap rows{10}, elems{12}, nodes{10};
};

using Leaf = std::pair<int, int>;  // rows order[first, second)

struct Partitioner {
    int N, nE, max_rows, budget, slot, block;
    bool ela;
    const int *e2n, *row, *col;
    const std::vector<int> &n2eI, &n2eV;
    const double *coord;
    std::vector<int> order;
    int spawn_depth = 0;
    std::atomic<int> failed{0};

    Partitioner(int N_, int nE_, int mr, int bud, int slot_, int block_, bool ela_,
                const int *e2n_, const int *row_, const int *col_, const std::vector<int> &I,
                const std::vector<int> &V, const double *c, std::vector<int> rows)
        : N(N_), nE(nE_), max_rows(mr), budget(bud), slot(slot_), block(block_), ela(ela_),
          e2n(e2n_), row(row_), col(col_), n2eI(I), n2eV(V), coord(c), order(std::move(rows))
    {
    }

    // Whether the tile made of rows order[lo, hi) fits the LDS budget and every
    // per-tile limit of the kernel (each limit is checked as soon as it can fail).
    bool fits(int lo, int hi, FitScratch &s) const
    {
        const int R = hi - lo;
        if (R > max_rows || R > kMaxTileRows) return false;
        s.rows.clear();
        for (int i = lo; i < hi; i++) s.rows.insert(order[i], 0);
        // tasks: at most one per off-diagonal block, a pair of the tile's rows once
        int64_t tasks = 0, ent = 0, inc = 0;
        for (int i = lo; i < hi; i++) {
            const int r = order[i];
            ent += row[r + 1] - row[r];
            for (int q = row[r]; q < row[r + 1]; q++) {
                const int c = col[q] - 1;
                if (c != r && (s.rows.get(c) < 0 || c > r)) tasks++;
            }
            inc += (n2eI[r + 1] - n2eI[r] + kIncAlign - 1) / kIncAlign * kIncAlign;
        }
        const int64_t inc8 = (inc + 7) / 8 * 8;
        if (inc8 > kMaxTileInc || ent > kMaxLdsBytes) return false;
        if (!tile_blocks_fit(ela, block, diag_blocks(ela, R), (int)((tasks + kLanes - 1) / kLanes)))
            return false;
        s.elems.clear();
        s.nodes.clear();
        for (int i = lo; i < hi; i++) s.nodes.insert(order[i], 0);
        int TE = 0, NN = R;
        const int maxTE = max_tile_elems(block);
        for (int i = lo; i < hi; i++) {
            const int r = order[i];
            for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                const int e = n2eV[j];
                if (s.elems.insert(e, 0) != 1) continue;
                if (++TE > maxTE) return false;
                for (int k = 0; k < 4; k++)
                    if (s.nodes.insert(e2n[(int64_t)e * 4 + k] - 1, 0) == 1 && ++NN > kMaxTileNodes)
                        return false;
            }
        }
        return tile_lds_layout(ela, NN, TE, TE, R, (int)inc8, (int)tasks, (int)ent, slot).total <=
               budget;
    }

    // Recursive coordinate bisection along the longest axis. A region whose halves do
    // not fit but whose thirds do is cut in three: tiles of a third rather than a quarter
    // of it, i.e. fewer and larger tiles. The top levels run their halves on two threads.
    // known: whether order[lo, hi) fits, when the caller has tested it (-1: not tested)
    void rcb(int lo, int hi, FitScratch &s, std::vector<Leaf> &out, int depth, int known = -1)
    {
        if (hi - lo <= 1 || known == 1 || (known < 0 && fits(lo, hi, s))) {
            if (hi > lo) out.emplace_back(lo, hi);
            return;
        }
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
        for (int i = lo; i < hi; i++)
            for (int d = 0; d < 3; d++) {
                const double x = coord[(int64_t)order[i] * 3 + d];
                mn[d] = std::min(mn[d], x);
                mx[d] = std::max(mx[d], x);
            }
        int ax = 0;
        for (int d = 1; d < 3; d++)
            if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
        const double *c = coord;
        auto less = [c, ax](int a, int b) {
            const double xa = c[(int64_t)a * 3 + ax], xb = c[(int64_t)b * 3 + ax];
            return xa < xb || (xa == xb && a < b);
        };
        // A cut near position p of order[a, b): the range is partitioned at p along the
        // axis, then the cut moves to the widest coordinate gap within w of p. On a layered
        // (e.g. jittered structured) mesh that gap lies between two layers of nodes, so a
        // tile's faces are flat and its halo thin: at 100^3, 1.71 instead of 1.97 staged
        // elements per element for lap tiles, which also grow larger under the same node cap.
        auto cut = [&](int a, int b, int p, int w) {
            std::nth_element(order.begin() + a, order.begin() + p, order.begin() + b, less);
            const int l = std::max(a + 1, p - w), r = std::min(b - 1, p + w);
            if (r - l < 2) return p;
            std::nth_element(order.begin() + a, order.begin() + l, order.begin() + p, less);
            std::nth_element(order.begin() + p, order.begin() + r, order.begin() + b, less);
            std::sort(order.begin() + l, order.begin() + r, less);
            double best = -1;
            int bi = p;
            for (int i = l + 1; i < r; i++) {
                const double g = c[(int64_t)order[i] * 3 + ax] - c[(int64_t)order[i - 1] * 3 + ax];
                if (g > best) {
                    best = g;
                    bi = i;
                }
            }
            return bi;
        };
        // whether both halves of a cut at m fit; fl/fr: each half's result (-1: not tested)
        auto halves_fit = [&](int m, int &fl, int &fr) {
            fl = fits(lo, m, s) ? 1 : 0;
            fr = fl ? (fits(m, hi, s) ? 1 : 0) : -1;
            return fl == 1 && fr == 1;
        };
        const int half = lo + (hi - lo) / 2;
        int mid = cut(lo, hi, half, (hi - lo) / 8), fl, fr;
        bool both = halves_fit(mid, fl, fr);
        if (!both) {
            // the flat cut leaves a piece too big: the median may still give two tiles
            // where the flat one would give more
            std::nth_element(order.begin() + lo, order.begin() + half, order.begin() + hi, less);
            int hl, hr;
            if (halves_fit(half, hl, hr)) {
                mid = half;
                fl = hl;
                fr = hr;
                both = true;
            } else {
                mid = cut(lo, hi, half, (hi - lo) / 8);  // the same halves as the first cut
            }
        }
        // thirds can fit only when the first one has at most max_rows rows
        if (hi - lo >= 6 && (hi - lo) / 3 <= max_rows && !both) {
            std::vector<int> saved(order.begin() + lo, order.begin() + hi);
            const int t1 = cut(lo, hi, lo + (hi - lo) / 3, (hi - lo) / 12);
            const int t2 = cut(t1, hi, lo + 2 * (hi - lo) / 3, (hi - lo) / 12);
            if (fits(lo, t1, s) && fits(t1, t2, s) && fits(t2, hi, s)) {
                out.emplace_back(lo, t1);
                out.emplace_back(t1, t2);
                out.emplace_back(t2, hi);
                return;
            }
            std::copy(saved.begin(), saved.end(), order.begin() + lo);  // back to the halves
        }
        // fl/fr hold for the halves of this cut: the recursion need not test them again
        if (depth < spawn_depth && hi - lo > 4 * max_rows) {
            std::vector<Leaf> left;
            bool spawned = false;
            std::thread t;
            try {
                t = std::thread([&] {
                    try {
                        FitScratch s2;
                        rcb(lo, mid, s2, left, depth + 1, fl);
                    } catch (...) {
                        failed = 1;
                    }
                });
                spawned = true;
            } catch (...) {
            }
            std::vector<Leaf> right;
            try {
                if (!spawned) rcb(lo, mid, s, left, depth + 1, fl);
                rcb(mid, hi, s, right, depth + 1, fr);
            } catch (...) {  // join the other half before anything leaves this frame
                failed = 1;
            }
            if (spawned) t.join();
            if (failed) return;
            out.insert(out.end(), left.begin(), left.end());
            out.insert(out.end(), right.begin(), right.end());
            return;
        }
        rcb(lo, mid, s, out, depth + 1, fl);
        rcb(mid, hi, s, out, depth + 1, fr);
    }

    // Breadth-first order of the node graph (through shared elements), then greedy
    // chunks that fit.
    void bfs_chunks(std::vector<Leaf> &out)
    {
        std::vector<char> isRow(N, 0), seen(N, 0);
        for (int r : order) isRow[r] = 1;
        std::vector<int> q;
        q.reserve(order.size());
        for (int s : std::vector<int>(order)) {
            if (seen[s]) continue;
            seen[s] = 1;
            size_t head = q.size();
            q.push_back(s);
            while (head < q.size()) {
                const int r = q[head++];
                for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                    const int e = n2eV[j];
                    for (int k = 0; k < 4; k++) {
                        const int v = e2n[(int64_t)e * 4 + k] - 1;
                        if (!seen[v] && isRow[v]) { seen[v] = 1; q.push_back(v); }
                    }
                }
            }
        }
        order = q;
        FitScratch s;
        const int n = (int)order.size();
        int lo = 0;
        while (lo < n) {
            int hi = lo + 1;
            while (hi < n && fits(lo, hi + 1, s)) hi++;
            out.emplace_back(lo, hi);
            lo = hi;
        }
    }
};

// The tiles of one contiguous range of leaves, with offsets into its own arrays.
struct TileChunk {
    std::vector<TileHdr> tiles;
    std::vector<int32_t> tile_elems;
    std::vector<uint8_t> tile_verts;
    std::vector<int32_t> tile_nodes;
    std::vector<RowInfo> tile_rowinfo;
    std::vector<uint16_t> inc;
    std::vector<uint32_t> task_rec;
    int max_node = 0, max_elem = 0, max_rows = 0;
    int status = MF_OK;
};

struct Emitter {
    bool ela;
    int N;
    const int *e2n, *row, *col, *checkBounds;
    const std::vector<int> &n2eI, &n2eV, &order, &nodeTile;
    const std::vector<Leaf> &leaves;

    int64_t find_pos(int c, int a) const  // CSR position of block (c, a)
    {
        for (int q = row[c]; q < row[c + 1]; q++)
            if (col[q] == a + 1) return q;
        return -1;
    }

    // Tile ti appended to o.
    // Per-row scratch of the task records: for each column of the row, the elements
    // holding both nodes (the first kListLen as contributions), their incidence mask and
    // count, filled in one pass over the row's incidence list.
    static constexpr int kMaxRowLen = 256;  // the task record's 8-bit column offsets
    struct ColScratch {
        SmallMap local{10};  // column node -> its position in the row
        uint16_t list[kMaxRowLen][kListLen];
        uint64_t mask[kMaxRowLen];
        int cnt[kMaxRowLen];
    };

    // Tile ti appended to o.
    int emit(int ti, TileChunk &o, SmallMap &nodeLocal, SmallMap &elemLocal,
             std::vector<int> &rows, std::vector<int> &tel, ColScratch &cs) const
    {
        const Leaf &lf = leaves[ti];
        rows.assign(order.begin() + lf.first, order.begin() + lf.second);
        std::sort(rows.begin(), rows.end());
        const int R = (int)rows.size();
        if (R > kMaxTileRows) return MF_ERR_ARG;
        TileHdr h{};
        h.node_start = (int)o.tile_nodes.size();
        h.n_rows = R;
        h.rinfo_start = (int)o.tile_rowinfo.size();
        nodeLocal.clear();
        for (int i = 0; i < R; i++) {
            nodeLocal.insert(rows[i], i);
            o.tile_nodes.push_back(rows[i]);
        }
        tel.clear();
        elemLocal.clear();
        for (int r : rows)
            for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                const int ins = elemLocal.insert(n2eV[j], 0);
                if (ins == -2) return MF_ERR_ARG;
                if (ins == 1) tel.push_back(n2eV[j]);
            }
        std::sort(tel.begin(), tel.end());
        const int TE = (int)tel.size();
        if (TE > kMaxTileElems) return MF_ERR_ARG;
        elemLocal.clear();
        for (int t = 0; t < TE; t++) elemLocal.insert(tel[t], t);
        for (int t = 0; t < TE; t++)
            for (int k = 0; k < 4; k++) {
                const int v = e2n[(int64_t)tel[t] * 4 + k] - 1;
                const int ins = nodeLocal.insert(v, (int)o.tile_nodes.size() - h.node_start);
                if (ins == -2) return MF_ERR_ARG;
                if (ins == 1) o.tile_nodes.push_back(v);
            }
        h.n_node = (int)o.tile_nodes.size() - h.node_start;
        if (h.n_node > kMaxTileNodes) return MF_ERR_ARG;
        h.elem_start = (int)o.tile_elems.size();
        h.n_elem = TE;
        for (int t = 0; t < TE; t++) {
            o.tile_elems.push_back(tel[t]);
            for (int k = 0; k < 4; k++)
                o.tile_verts.push_back((uint8_t)nodeLocal.get(e2n[(int64_t)tel[t] * 4 + k] - 1));
        }

        // tile entries: the rows' CSR segments in tile row order
        std::vector<int> ent((size_t)R + 1, 0);
        for (int ra = 0; ra < R; ra++) ent[ra + 1] = ent[ra] + row[rows[ra] + 1] - row[rows[ra]];
        h.n_ent = ent[R];

        // incidence lists (ascending element = ascending tile element) and entries
        h.inc_start = (int)o.inc.size();
        h.task_start = (int)(o.task_rec.size() / 4);
        int incOff = 0;
        for (int ra = 0; ra < R; ra++) {
            const int a = rows[ra], len = row[a + 1] - row[a];
            const int val = n2eI[a + 1] - n2eI[a];
            int doff = -1;
            for (int q = 0; q < len; q++)
                if (col[row[a] + q] - 1 == a) { doff = q; break; }
            if (doff < 0 || doff > 0xFF) return MF_ERR_ARG;
            int mask = 0;
            if (ela && checkBounds)
                for (int k = 0; k < 3; k++)
                    if (checkBounds[(int64_t)k * N + a] != 0) mask |= 1 << k;
            o.tile_rowinfo.push_back(RowInfo{row[a], len | doff << 16 | mask << 24, incOff | val << 16, ent[ra]});
            for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                const int e = n2eV[j];
                int va = 0;
                for (int k = 0; k < 4; k++)
                    if (e2n[(int64_t)e * 4 + k] - 1 == a) va = k;
                o.inc.push_back((uint16_t)(elemLocal.get(e) << 2 | va));
            }
            // the elements holding a and each column c, ascending: one pass over a's
            // incidence list, each element filed under the columns of its other vertices
            // (a repeated vertex counts once, with its last slot, for a and for c)
            if (len > kMaxRowLen) return MF_ERR_ARG;
            cs.local.clear();
            for (int q = 0; q < len; q++) {
                cs.local.insert(col[row[a] + q] - 1, q);
                cs.cnt[q] = 0;
                cs.mask[q] = 0;
            }
            for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                const int *v = e2n + (int64_t)n2eV[j] * 4;
                int va = -1;
                for (int k = 0; k < 4; k++)
                    if (v[k] - 1 == a) va = k;
                const int tl = elemLocal.get(n2eV[j]);
                for (int k = 0; k < 4; k++) {
                    const int c = v[k] - 1;
                    if (c == a) continue;
                    bool later = false;
                    for (int k2 = k + 1; k2 < 4; k2++) later |= v[k2] == v[k];
                    if (later) continue;
                    const int q = cs.local.get(c);
                    if (q < 0) return MF_ERR_ARG;  // a vertex outside the row's pattern
                    if (j - n2eI[a] < kMaskValence) cs.mask[q] |= 1ull << (j - n2eI[a]);
                    if (cs.cnt[q] < kListLen) cs.list[q][cs.cnt[q]] = (uint16_t)(tl << 4 | va << 2 | k);
                    cs.cnt[q]++;
                }
            }
            for (int q = 0; q < len; q++) {
                const int c = col[row[a] + q] - 1;
                if (c < 0 || c >= N || nodeLocal.get(c) < 0 || q > 0xFF) return MF_ERR_ARG;
                if (c == a) continue;
                // pair (a,c)/(c,a) inside one tile: computed once, by the row with the
                // smaller node id, the transpose written into the tile's output image;
                // a column of another tile: that tile computes (c,a) itself
                int oc = 0;
                if (nodeTile[c] == ti) {
                    if (c < a) continue;
                    const int64_t pc = find_pos(c, a);
                    if (pc < 0) return MF_ERR_ARG;  // the pattern is not symmetric
                    oc = (int)(pc - row[c]);
                    if (oc > 0xFF) return MF_ERR_ARG;
                }
                const int cnt = cs.cnt[q];
                if (cnt < 1) return MF_ERR_ARG;
                if (cnt <= kListLen) {
                    uint16_t list[kListLen];
                    for (int z = 0; z < kListLen; z++) list[z] = z < cnt ? cs.list[q][z] : (uint16_t)(TE << 4);
                    for (int w = 0; w < 3; w++)
                        o.task_rec.push_back((uint32_t)list[2 * w] | (uint32_t)list[2 * w + 1] << 16);
                } else if (val <= kMaskValence) {
                    o.task_rec.push_back(kMaskRecord);
                    o.task_rec.push_back((uint32_t)cs.mask[q]);
                    o.task_rec.push_back((uint32_t)(cs.mask[q] >> 32));
                } else {
                    o.task_rec.push_back(kWalkRecord);
                    o.task_rec.push_back(0u);
                    o.task_rec.push_back(0u);
                }
                o.task_rec.push_back((uint32_t)(ra | q << 8 | nodeLocal.get(c) << 16 | oc << 24));
            }
            while (o.inc.size() % kIncAlign) o.inc.push_back(0);
            incOff = (int)o.inc.size() - h.inc_start;
        }
        while (o.inc.size() % 8) o.inc.push_back(0);
        h.n_inc = (int)o.inc.size() - h.inc_start;
        h.n_task = (int)(o.task_rec.size() / 4) - h.task_start;
        o.tiles.push_back(h);
        o.max_node = std::max(o.max_node, h.n_node);
        o.max_elem = std::max(o.max_elem, TE);
        o.max_rows = std::max(o.max_rows, R);
        if (o.inc.size() > (size_t)INT32_MAX - 100000 || o.task_rec.size() > (size_t)INT32_MAX / 2)
            return MF_ERR_NOMEM;
        return MF_OK;
    }

    void emit_range(int t0, int t1, TileChunk &o) const
    {
        try {
            SmallMap nodeLocal(10), elemLocal(12);
            std::vector<int> rows, tel;
            auto cs = std::make_unique<ColScratch>();
            for (int ti = t0; ti < t1 && o.status == MF_OK; ti++)
                o.status = emit(ti, o, nodeLocal, elemLocal, rows, tel, *cs);
        } catch (...) {
            o.status = MF_ERR_NOMEM;
        }
    }
};

}  // namespace

int build_plan(PlanHost &p, int op, int nE, int N, int nnz, const int *e2n, const int *row,
               const int *col, const int *checkBounds, const double *coord, int lds_budget,
               int max_rows, int slot_bytes, int block, int id_chunks)
{
    p = PlanHost{};
    p.op = op;
    p.nE = nE;
    p.N = N;
    p.nnz = nnz;
    p.block = block;
    const bool ela = op == MF_OP_ELA;
    if (nE < 0 || N < 0 || nnz < 0 || max_rows < 1 || max_rows > kMaxTileRows || id_chunks < 1)
        return MF_ERR_ARG;
    if (block != 512 && block != 1024) return MF_ERR_ARG;
    if (row[0] != 0 || row[N] != nnz) return MF_ERR_ARG;
    std::vector<int> rowsAll;
    for (int i = 0; i < N; i++) {
        const int len = row[i + 1] - row[i];
        if (len < 0 || len > 0xFFFF) return MF_ERR_ARG;
        if (len == 0) p.isolated.push_back(i);
        else rowsAll.push_back(i);
    }
    // each node's elements, ascending (the threaded counting fill of the setup)
    std::vector<int> n2eI((size_t)N + 1, 0), n2eV((size_t)nE * 4 + 1);
    const int st = mf_create_nodeToElem(e2n, nE, N, n2eI.data(), n2eV.data());
    if (st != MF_OK) return st;
    int maxVal = 0;
    for (int i = 0; i < N; i++) maxVal = std::max(maxVal, n2eI[i + 1] - n2eI[i]);
    if (maxVal > kMaxValence) return MF_ERR_ARG;
    // a node with elements has a non-empty row holding its diagonal
    for (int i = 0; i < N; i++)
        if ((n2eI[i + 1] > n2eI[i]) != (row[i + 1] > row[i])) return MF_ERR_ARG;

    const int threads = plan_threads();
    Partitioner part(N, nE, max_rows, lds_budget, slot_bytes, block, ela, e2n, row, col, n2eI,
                     n2eV, coord, rowsAll);
    while ((1 << part.spawn_depth) < threads) part.spawn_depth++;
    std::vector<Leaf> leaves;
    if (coord) {
        // rowsAll is in ascending node id: contiguous ranges of it are id ranges
        FitScratch s;
        const int n = (int)rowsAll.size(), K = std::min(id_chunks, std::max(n, 1));
        for (int k = 0; k < K; k++)
            part.rcb((int)((int64_t)n * k / K), (int)((int64_t)n * (k + 1) / K), s, leaves, 0);
    } else {
        part.bfs_chunks(leaves);
    }
    if (part.failed) return MF_ERR_NOMEM;

    // tile of every row node (the kernel runs tile i on XCD i / ceil(T / 8))
    const int T = (int)leaves.size();
    std::vector<int> nodeTile((size_t)N, -1);
    for (int i = 0; i < T; i++)
        for (int q = leaves[i].first; q < leaves[i].second; q++) nodeTile[part.order[q]] = i;

    // the tiles in contiguous chunks, one per thread, then joined in order
    const Emitter em{ela, N, e2n, row, col, checkBounds, n2eI, n2eV, part.order, nodeTile, leaves};
    const int nc = std::max(1, std::min(threads, T / 64));
    std::vector<TileChunk> chunks((size_t)nc);
    std::vector<std::thread> pool;
    pool.reserve((size_t)nc);
    for (int c = 1; c < nc; c++) {
        try {
            pool.emplace_back([&, c] { em.emit_range((int)((int64_t)T * c / nc),
                                                     (int)((int64_t)T * (c + 1) / nc), chunks[c]); });
        } catch (...) {  // no thread: run that chunk here
            em.emit_range((int)((int64_t)T * c / nc), (int)((int64_t)T * (c + 1) / nc), chunks[c]);
        }
    }
    em.emit_range(0, (int)((int64_t)T / nc), chunks[0]);  // catches its own exceptions
    for (auto &t : pool) t.join();
    size_t nTiles = 0, nElems = 0, nNodes = 0, nRinfo = 0, nInc = 0, nTask = 0;
    for (const TileChunk &c : chunks) {
        if (c.status != MF_OK) return c.status;
        nTiles += c.tiles.size();
        nElems += c.tile_elems.size();
        nNodes += c.tile_nodes.size();
        nRinfo += c.tile_rowinfo.size();
        nInc += c.inc.size();
        nTask += c.task_rec.size();
    }
    if (nInc > (size_t)INT32_MAX - 100000 || nTask > (size_t)INT32_MAX / 2 ||
        nNodes > (size_t)INT32_MAX || nElems > (size_t)INT32_MAX / 4)
        return MF_ERR_NOMEM;
    p.tiles.reserve(nTiles);
    p.tile_elems.reserve(nElems);
    p.tile_verts.reserve(nElems * 4);
    p.tile_nodes.reserve(nNodes);
    p.tile_rowinfo.reserve(nRinfo);
    p.inc.reserve(nInc);
    p.task_rec.reserve(nTask);
    for (TileChunk &c : chunks) {
        const int oN = (int)p.tile_nodes.size(), oE = (int)p.tile_elems.size(),
                  oR = (int)p.tile_rowinfo.size(), oI = (int)p.inc.size(),
                  oT = (int)(p.task_rec.size() / 4);
        for (TileHdr h : c.tiles) {
            h.node_start += oN;
            h.elem_start += oE;
            h.rinfo_start += oR;
            h.inc_start += oI;
            h.task_start += oT;
            p.tiles.push_back(h);
        }
        p.tile_elems.insert(p.tile_elems.end(), c.tile_elems.begin(), c.tile_elems.end());
        p.tile_verts.insert(p.tile_verts.end(), c.tile_verts.begin(), c.tile_verts.end());
        p.tile_nodes.insert(p.tile_nodes.end(), c.tile_nodes.begin(), c.tile_nodes.end());
        p.tile_rowinfo.insert(p.tile_rowinfo.end(), c.tile_rowinfo.begin(), c.tile_rowinfo.end());
        p.inc.insert(p.inc.end(), c.inc.begin(), c.inc.end());
        p.task_rec.insert(p.task_rec.end(), c.task_rec.begin(), c.task_rec.end());
        p.max_node = std::max(p.max_node, c.max_node);
        p.max_elem = std::max(p.max_elem, c.max_elem);
        p.max_rows = std::max(p.max_rows, c.max_rows);
        c = TileChunk{};  // release as we go
    }
    return MF_OK;
}

namespace {

// Which positions of a row's incidence list (up to kMaxValence) hold a tile node.
using Holds = unsigned __int128;
static_assert(kMaxValence <= 128, "Holds has one bit per incidence entry");
inline int popcount128(Holds x)
{
    return __builtin_popcountll((uint64_t)x) + __builtin_popcountll((uint64_t)(x >> 64));
}

// Claims a CSR entry or a row for one writer; false when it was claimed before.
inline bool claim(uint8_t *flag) { return __atomic_fetch_add(flag, 1, __ATOMIC_RELAXED) == 0; }

struct TileValidator {
    const PlanHost &p;
    const int *row, *col, *e2n;
    int slot_bytes;
    bool ela;
    size_t nTasks;
    uint8_t *hit, *rowSeen;  // shared by all tiles, claimed atomically

    int check(const TileHdr &h, SmallMap &localOf, std::vector<int> &cnt,
              std::vector<Holds> &holds) const
    {
        const int R = h.n_rows;
        if (R < 1 || R > kMaxTileRows || h.n_node < R || h.n_node > kMaxTileNodes ||
            h.n_elem < 0 || h.n_elem > max_tile_elems(p.block))
            return MF_ERR_STATE;
        if (h.node_start < 0 || (size_t)h.node_start + h.n_node > p.tile_nodes.size()) return MF_ERR_STATE;
        if (h.elem_start < 0 || (size_t)h.elem_start + h.n_elem > p.tile_elems.size()) return MF_ERR_STATE;
        if (h.rinfo_start < 0 || (size_t)h.rinfo_start + R > p.tile_rowinfo.size()) return MF_ERR_STATE;
        if (h.inc_start < 0 || h.inc_start % 8 || h.n_inc < 0 || h.n_inc % 8 ||
            h.n_inc > kMaxTileInc || (size_t)h.inc_start + h.n_inc > p.inc.size())
            return MF_ERR_STATE;
        if (h.task_start < 0 || h.n_task < 0 || (size_t)h.task_start + h.n_task > nTasks)
            return MF_ERR_STATE;
        // one work block per compute wave, and a compute wave without a diagonal block
        const int nDB = diag_blocks(ela, R), nEB = (h.n_task + kLanes - 1) / kLanes;
        if (!tile_blocks_fit(ela, p.block, nDB, nEB)) return MF_ERR_STATE;
        if (h.n_ent < R || h.n_node > p.max_node || h.n_elem > p.max_elem ||
            tile_lds_layout(ela, p.max_node, p.max_elem, h.n_elem, R, h.n_inc, h.n_task, h.n_ent,
                            slot_bytes).total + 16 > kMaxLdsBytes)
            return MF_ERR_STATE;
        // row entry offsets: consecutive segments in tile row order
        {
            int e = 0;
            for (int ra = 0; ra < R; ra++) {
                const RowInfo ri = p.tile_rowinfo[h.rinfo_start + ra];
                if (ri.ent != e) return MF_ERR_STATE;
                e += ri.len_diag & 0xFFFF;
            }
            if (e != h.n_ent) return MF_ERR_STATE;
        }
        const int32_t *nodes = p.tile_nodes.data() + h.node_start;
        const uint8_t *verts = p.tile_verts.data() + (size_t)h.elem_start * 4;
        localOf.clear();
        for (int i = 0; i < h.n_node; i++) {
            const int n = nodes[i];
            if (n < 0 || n >= p.N || localOf.insert(n, i) != 1) return MF_ERR_STATE;
            if (i < R && !claim(rowSeen + n)) return MF_ERR_STATE;
        }
        for (int t = 0; t < h.n_elem; t++) {
            const int e = p.tile_elems[h.elem_start + t];
            if (e < 0 || e >= p.nE) return MF_ERR_STATE;
            if (t > 0 && e <= p.tile_elems[h.elem_start + t - 1]) return MF_ERR_STATE;
            for (int k = 0; k < 4; k++)
                if (verts[t * 4 + k] >= h.n_node || nodes[verts[t * 4 + k]] != e2n[(int64_t)e * 4 + k] - 1)
                    return MF_ERR_STATE;
        }
        for (int i = h.inc_start; i < h.inc_start + h.n_inc; i++)
            if ((p.inc[i] >> 2) >= std::max(h.n_elem, 1)) return MF_ERR_STATE;  // padding too
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int q = 0; q < h.n_elem * 4; q++) cnt[verts[q]]++;
        int incNext = 0, tk = h.task_start;
        const int tkEnd = h.task_start + h.n_task;
        for (int ra = 0; ra < R; ra++) {
            const int a = nodes[ra];
            const RowInfo ri = p.tile_rowinfo[h.rinfo_start + ra];
            const int len = ri.len_diag & 0xFFFF, doff = (ri.len_diag >> 16) & 0xFF;
            const int mask = (ri.len_diag >> 24) & 0xFF;
            const int incOff = ri.inc & 0xFFFF, val = (ri.inc >> 16) & 0xFFFF;
            if (ri.rowptr != row[a] || len != row[a + 1] - row[a] || len < 1 || doff >= len ||
                col[row[a] + doff] != a + 1 || mask > 7 || (!ela && mask))
                return MF_ERR_STATE;
            if (incOff != incNext || incOff % kIncAlign || val < 1 || val > kMaxValence ||
                incOff + val > h.n_inc)
                return MF_ERR_STATE;
            incNext += (val + kIncAlign - 1) / kIncAlign * kIncAlign;
            if (!claim(hit + row[a] + doff)) return MF_ERR_STATE;  // the diagonal block
            // the incidence list: every element of the tile holding a, ascending
            int last = -1;
            for (int i = 0; i < val; i++) {
                const uint16_t en = p.inc[(size_t)h.inc_start + incOff + i];
                const int t = en >> 2, va = en & 3;
                if (t >= h.n_elem || t <= last || verts[t * 4 + va] != ra) return MF_ERR_STATE;
                last = t;
                for (int k = 0; k < 4; k++) holds[verts[t * 4 + k]] |= (Holds)1 << i;
            }
            if (cnt[ra] != val) return MF_ERR_STATE;
            // the row's tasks: exact masks, each block and its transpose written once
            for (; tk < tkEnd && (int)(p.task_rec[(size_t)tk * 4 + 3] & 0xFF) == ra; tk++) {
                const uint32_t *rec = p.task_rec.data() + (size_t)tk * 4;
                const int oa = (rec[3] >> 8) & 0xFF, cl = (rec[3] >> 16) & 0xFF, oc = rec[3] >> 24;
                if (oa >= len) return MF_ERR_STATE;
                const int c = col[row[a] + oa] - 1;
                if (c == a || !claim(hit + row[a] + oa)) return MF_ERR_STATE;
                const int ec = localOf.get(c);
                const int nHold = ec < 0 ? 0 : popcount128(holds[ec]);
                if (ec < 0 || cl != ec || nHold < 1) return MF_ERR_STATE;
                if (nHold <= kListLen) {
                    // the exact ascending list of (element, slot of a, slot of c), padded
                    // with the zero slot
                    int q = 0;
                    for (int i = 0; i < val; i++) {
                        if (!((holds[ec] >> i) & 1)) continue;
                        const uint16_t en = p.inc[(size_t)h.inc_start + incOff + i];
                        const int t = en >> 2;
                        int vc = -1;
                        for (int k = 0; k < 4; k++)
                            if (verts[t * 4 + k] == ec) vc = k;
                        const uint16_t want = (uint16_t)(t << 4 | (en & 3) << 2 | vc);
                        if (((rec[q / 2] >> (16 * (q % 2))) & 0xFFFF) != want) return MF_ERR_STATE;
                        q++;
                    }
                    for (; q < kListLen; q++)
                        if (((rec[q / 2] >> (16 * (q % 2))) & 0xFFFF) != (uint32_t)(h.n_elem << 4))
                            return MF_ERR_STATE;
                } else if (val <= kMaskValence) {
                    const uint64_t m = (uint64_t)rec[1] | (uint64_t)rec[2] << 32;
                    if (rec[0] != kMaskRecord || (Holds)m != holds[ec]) return MF_ERR_STATE;
                } else if (rec[0] != kWalkRecord || rec[1] != 0 || rec[2] != 0) {
                    return MF_ERR_STATE;  // the kernel finds the elements on row a's list
                }
                if (ec < R) {  // c is a row of this tile: the task also writes (c, a)
                    if (oc >= (p.tile_rowinfo[h.rinfo_start + ec].len_diag & 0xFFFF) ||
                        col[row[c] + oc] != a + 1 || !claim(hit + row[c] + oc))
                        return MF_ERR_STATE;
                } else if (oc != 0) {
                    return MF_ERR_STATE;
                }
            }
            for (int i = 0; i < val; i++) {
                const int t = p.inc[(size_t)h.inc_start + incOff + i] >> 2;
                for (int k = 0; k < 4; k++) holds[verts[t * 4 + k]] = 0;
            }
        }
        if (tk != tkEnd || incNext > h.n_inc || h.n_inc - incNext >= 8) return MF_ERR_STATE;
        return MF_OK;
    }
};

}  // namespace

int validate_plan(const PlanHost &p, const int *row, const int *col, const int *e2n,
                  int slot_bytes)
{
    if (p.tile_verts.size() != p.tile_elems.size() * 4) return MF_ERR_STATE;
    if (p.task_rec.size() % 4) return MF_ERR_STATE;
    std::vector<uint8_t> hit((size_t)p.nnz, 0), rowSeen((size_t)p.N, 0);
    for (int i : p.isolated) {
        if (i < 0 || i >= p.N || row[i + 1] != row[i] || rowSeen[i]) return MF_ERR_STATE;
        rowSeen[i] = 1;
    }
    const TileValidator tv{p, row, col, e2n, slot_bytes, p.op == MF_OP_ELA, p.task_rec.size() / 4,
                           hit.data(), rowSeen.data()};
    // the tiles in contiguous chunks, one per thread; the first failure in any stops all
    const int T = (int)p.tiles.size();
    const int nc = std::max(1, std::min(plan_threads(), T / 64));
    std::vector<int> status((size_t)nc, MF_OK);
    std::atomic<int> stop{0};
    auto run = [&](int c) {
        try {
            SmallMap localOf(10);
            std::vector<int> cnt(kMaxTileNodes);
            std::vector<Holds> holds(kMaxTileNodes, 0);
            const int t1 = (int)((int64_t)T * (c + 1) / nc);
            for (int t = (int)((int64_t)T * c / nc); t < t1 && !stop; t++) {
                const int st = tv.check(p.tiles[t], localOf, cnt, holds);
                if (st != MF_OK) {
                    status[c] = st;
                    stop = 1;
                }
            }
        } catch (...) {
            status[c] = MF_ERR_NOMEM;
            stop = 1;
        }
    };
    std::vector<std::thread> pool;
    pool.reserve((size_t)nc);
    for (int c = 1; c < nc; c++) {
        try {
            pool.emplace_back(run, c);
        } catch (...) {
            run(c);
        }
    }
    run(0);
    for (auto &t : pool) t.join();
    for (int st : status)
        if (st != MF_OK) return st;
    for (int i = 0; i < p.N; i++)
        if (rowSeen[i] != 1) return MF_ERR_STATE;
    for (int64_t q = 0; q < p.nnz; q++)
        if (hit[q] != 1) return MF_ERR_STATE;
    return MF_OK;
}

}  // namespace mf
