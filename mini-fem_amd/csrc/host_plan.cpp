// Host construction and validation of the tiled gather schedule (layout in plan.h).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "minifem.h"
#include "plan.h"

namespace mf {

namespace {

struct Partitioner {
    int N, nE, max_rows, budget, slot, block;
    bool ela;
    const int *e2n, *row, *col;
    const std::vector<int> &n2eI, &n2eV;
    const double *coord;
    std::vector<int> order;
    std::vector<std::pair<int, int>> leaves;
    std::vector<int> elemStamp, nodeStamp;
    int stampId = 0;

    Partitioner(int N_, int nE_, int mr, int bud, int slot_, int block_, bool ela_,
                const int *e2n_, const int *row_, const int *col_, const std::vector<int> &I,
                const std::vector<int> &V, const double *c, std::vector<int> rows)
        : N(N_), nE(nE_), max_rows(mr), budget(bud), slot(slot_), block(block_), ela(ela_),
          e2n(e2n_), row(row_), col(col_), n2eI(I), n2eV(V), coord(c), order(std::move(rows)),
          elemStamp(nE_, -1), nodeStamp(N_, -1)
    {
    }

    // Whether the tile made of rows order[lo, hi) fits the LDS budget and every
    // per-tile limit of the kernel.
    bool fits(int lo, int hi)
    {
        const int R = hi - lo;
        if (R > max_rows || R > kMaxTileRows) return false;
        stampId++;
        for (int i = lo; i < hi; i++) nodeStamp[order[i]] = stampId;
        int64_t TE = 0, NN = R, inc = 0, tasks = 0, ent = 0;
        // tasks: at most one per off-diagonal block, a pair of the tile's rows once
        for (int i = lo; i < hi; i++) {
            const int r = order[i];
            ent += row[r + 1] - row[r];
            for (int q = row[r]; q < row[r + 1]; q++) {
                const int c = col[q] - 1;
                if (c != r && (nodeStamp[c] != stampId || c > r)) tasks++;
            }
        }
        for (int i = lo; i < hi; i++) {
            const int r = order[i];
            inc += (n2eI[r + 1] - n2eI[r] + kIncAlign - 1) / kIncAlign * kIncAlign;
            for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                const int e = n2eV[j];
                if (elemStamp[e] == stampId) continue;
                elemStamp[e] = stampId;
                TE++;
                for (int k = 0; k < 4; k++) {
                    const int v = e2n[(int64_t)e * 4 + k] - 1;
                    if (nodeStamp[v] != stampId) { nodeStamp[v] = stampId; NN++; }
                }
            }
        }
        if (NN > kMaxTileNodes || TE > max_tile_elems(block)) return false;
        const int64_t inc8 = (inc + 7) / 8 * 8;
        if (inc8 > kMaxTileInc || ent > kMaxLdsBytes) return false;
        if (!tile_blocks_fit(ela, block, diag_blocks(ela, R), (int)((tasks + kLanes - 1) / kLanes)))
            return false;
        return tile_lds_layout(ela, (int)NN, (int)TE, (int)TE, R, (int)inc8, (int)tasks, (int)ent,
                               slot).total <= budget;
    }

    // Recursive coordinate bisection along the longest axis. A region whose halves do
    // not fit but whose thirds do is cut in three: tiles of a third rather than a quarter
    // of it, i.e. fewer and larger tiles.
    void rcb(int lo, int hi)
    {
        if (hi - lo <= 1 || fits(lo, hi)) {
            if (hi > lo) leaves.emplace_back(lo, hi);
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
        const int mid = lo + (hi - lo) / 2;
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi, less);
        if (hi - lo >= 6 && !(fits(lo, mid) && fits(mid, hi))) {
            const int t1 = lo + (hi - lo) / 3, t2 = lo + 2 * (hi - lo) / 3;
            std::vector<int> saved(order.begin() + lo, order.begin() + hi);
            std::nth_element(order.begin() + lo, order.begin() + t1, order.begin() + hi, less);
            std::nth_element(order.begin() + t1, order.begin() + t2, order.begin() + hi, less);
            if (fits(lo, t1) && fits(t1, t2) && fits(t2, hi)) {
                leaves.emplace_back(lo, t1);
                leaves.emplace_back(t1, t2);
                leaves.emplace_back(t2, hi);
                return;
            }
            std::copy(saved.begin(), saved.end(), order.begin() + lo);  // back to the halves
        }
        rcb(lo, mid);
        rcb(mid, hi);
    }

    // Breadth-first order of the node graph (through shared elements), then greedy
    // chunks that fit.
    void bfs_chunks()
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
        const int n = (int)order.size();
        int lo = 0;
        while (lo < n) {
            int hi = lo + 1;
            while (hi < n && fits(lo, hi + 1)) hi++;
            leaves.emplace_back(lo, hi);
            lo = hi;
        }
    }
};

}  // namespace

int build_plan(PlanHost &p, int op, int nE, int N, int nnz, const int *e2n, const int *row,
               const int *col, const int *checkBounds, const double *coord, int lds_budget,
               int max_rows, int slot_bytes, int block)
{
    p = PlanHost{};
    p.op = op;
    p.nE = nE;
    p.N = N;
    p.nnz = nnz;
    p.block = block;
    const bool ela = op == MF_OP_ELA;
    if (nE < 0 || N < 0 || nnz < 0 || max_rows < 1 || max_rows > kMaxTileRows) return MF_ERR_ARG;
    if (block != 512 && block != 1024) return MF_ERR_ARG;
    if (row[0] != 0 || row[N] != nnz) return MF_ERR_ARG;
    std::vector<int> rowsAll;
    for (int i = 0; i < N; i++) {
        const int len = row[i + 1] - row[i];
        if (len < 0 || len > 0xFFFF) return MF_ERR_ARG;
        if (len == 0) p.isolated.push_back(i);
        else rowsAll.push_back(i);
    }
    std::vector<int> n2eI((size_t)N + 1, 0), n2eV((size_t)nE * 4 + 1);
    for (int64_t q = 0; q < (int64_t)nE * 4; q++) {
        const int v = e2n[q] - 1;
        if (v < 0 || v >= N) return MF_ERR_ARG;
        n2eI[v + 1]++;
    }
    int maxVal = 0;
    for (int i = 0; i < N; i++) {
        maxVal = std::max(maxVal, n2eI[i + 1]);
        n2eI[i + 1] += n2eI[i];
    }
    if (maxVal > kMaxValence) return MF_ERR_ARG;
    {
        std::vector<int> cur(n2eI.begin(), n2eI.end() - 1);
        for (int e = 0; e < nE; e++)
            for (int k = 0; k < 4; k++) n2eV[cur[e2n[(int64_t)e * 4 + k] - 1]++] = e;
    }
    // a node with elements has a non-empty row holding its diagonal
    for (int i = 0; i < N; i++)
        if ((n2eI[i + 1] > n2eI[i]) != (row[i + 1] > row[i])) return MF_ERR_ARG;

    Partitioner part(N, nE, max_rows, lds_budget, slot_bytes, block, ela, e2n, row, col, n2eI,
                     n2eV, coord, rowsAll);
    if (coord) part.rcb(0, (int)rowsAll.size());
    else part.bfs_chunks();

    // tile of every row node (the kernel runs tile i on XCD i / ceil(T / 8))
    const int T = (int)part.leaves.size();
    std::vector<int> nodeTile((size_t)N, -1);
    for (int i = 0; i < T; i++)
        for (int q = part.leaves[i].first; q < part.leaves[i].second; q++) nodeTile[part.order[q]] = i;
    auto find_pos = [&](int c, int a) -> int64_t {  // CSR position of block (c, a)
        for (int q = row[c]; q < row[c + 1]; q++)
            if (col[q] == a + 1) return q;
        return -1;
    };

    std::vector<int> nodeLocal((size_t)N, -1), elemLocal((size_t)nE, -1);
    std::vector<int> rows, tel;
    for (int ti = 0; ti < T; ti++) {
        const auto &lf = part.leaves[ti];
        rows.assign(part.order.begin() + lf.first, part.order.begin() + lf.second);
        std::sort(rows.begin(), rows.end());
        const int R = (int)rows.size();
        if (R > kMaxTileRows) return MF_ERR_ARG;
        TileHdr h{};
        h.node_start = (int)p.tile_nodes.size();
        h.n_rows = R;
        h.rinfo_start = (int)p.tile_rowinfo.size();
        for (int i = 0; i < R; i++) {
            nodeLocal[rows[i]] = i;
            p.tile_nodes.push_back(rows[i]);
        }
        tel.clear();
        for (int r : rows)
            for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                const int e = n2eV[j];
                if (elemLocal[e] == -1) { elemLocal[e] = -2; tel.push_back(e); }
            }
        std::sort(tel.begin(), tel.end());
        const int TE = (int)tel.size();
        if (TE > kMaxTileElems) return MF_ERR_ARG;
        for (int t = 0; t < TE; t++) elemLocal[tel[t]] = t;
        for (int t = 0; t < TE; t++)
            for (int k = 0; k < 4; k++) {
                const int v = e2n[(int64_t)tel[t] * 4 + k] - 1;
                if (nodeLocal[v] == -1) {
                    nodeLocal[v] = (int)p.tile_nodes.size() - h.node_start;
                    p.tile_nodes.push_back(v);
                }
            }
        h.n_node = (int)p.tile_nodes.size() - h.node_start;
        if (h.n_node > kMaxTileNodes) return MF_ERR_ARG;
        h.elem_start = (int)p.tile_elems.size();
        h.n_elem = TE;
        for (int t = 0; t < TE; t++) {
            p.tile_elems.push_back(tel[t]);
            for (int k = 0; k < 4; k++)
                p.tile_verts.push_back((uint8_t)nodeLocal[e2n[(int64_t)tel[t] * 4 + k] - 1]);
        }

        // tile entries: the rows' CSR segments in tile row order
        std::vector<int> ent((size_t)R + 1, 0);
        for (int ra = 0; ra < R; ra++) ent[ra + 1] = ent[ra] + row[rows[ra] + 1] - row[rows[ra]];
        h.n_ent = ent[R];

        // incidence lists (ascending element = ascending tile element) and entries
        h.inc_start = (int)p.inc.size();
        h.task_start = (int)(p.task_rec.size() / 4);
        int incOff = 0;
        for (int ra = 0; ra < R; ra++) {
            const int a = rows[ra], len = row[a + 1] - row[a];
            const int val = n2eI[a + 1] - n2eI[a];
            int doff = -1;
            for (int o = 0; o < len; o++)
                if (col[row[a] + o] - 1 == a) { doff = o; break; }
            if (doff < 0 || doff > 0xFF) return MF_ERR_ARG;
            int mask = 0;
            if (ela && checkBounds)
                for (int k = 0; k < 3; k++)
                    if (checkBounds[(int64_t)k * N + a] != 0) mask |= 1 << k;
            p.tile_rowinfo.push_back(RowInfo{row[a], len | doff << 16 | mask << 24, incOff | val << 16, ent[ra]});
            for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                const int e = n2eV[j];
                int va = 0;
                for (int k = 0; k < 4; k++)
                    if (e2n[(int64_t)e * 4 + k] - 1 == a) va = k;
                p.inc.push_back((uint16_t)(elemLocal[e] << 2 | va));
            }
            for (int o = 0; o < len; o++) {
                const int c = col[row[a] + o] - 1;
                if (c < 0 || c >= N || nodeLocal[c] < 0 || o > 0xFF) return MF_ERR_ARG;
                if (c == a) continue;
                // pair (a,c)/(c,a) inside one tile: computed once, by the row with the
                // smaller node id, the transpose written into the tile's output image;
                // a column of another tile: that tile computes (c,a) itself
                int oc = 0;
                if (nodeTile[c] == ti) {
                    if (c < a) continue;
                    const int64_t q = find_pos(c, a);
                    if (q < 0) return MF_ERR_ARG;  // the pattern is not symmetric
                    oc = (int)(q - row[c]);
                    if (oc > 0xFF) return MF_ERR_ARG;
                }
                // the elements holding a and c, ascending, with a's and c's slots
                uint64_t m = 0;
                uint16_t list[kListLen];
                int cnt = 0;
                for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                    const int *v = e2n + (int64_t)n2eV[j] * 4;
                    int va = -1, vc = -1;
                    for (int k = 0; k < 4; k++) {
                        if (v[k] - 1 == a) va = k;
                        if (v[k] - 1 == c) vc = k;
                    }
                    if (vc < 0) continue;
                    m |= 1ull << (j - n2eI[a]);
                    if (cnt < kListLen) list[cnt] = (uint16_t)(elemLocal[n2eV[j]] << 4 | va << 2 | vc);
                    cnt++;
                }
                if (cnt < 1) return MF_ERR_ARG;
                if (cnt <= kListLen) {
                    for (int q = cnt; q < kListLen; q++) list[q] = (uint16_t)(TE << 4);
                    for (int w = 0; w < 3; w++)
                        p.task_rec.push_back((uint32_t)list[2 * w] | (uint32_t)list[2 * w + 1] << 16);
                } else {
                    p.task_rec.push_back(kMaskRecord);
                    p.task_rec.push_back((uint32_t)m);
                    p.task_rec.push_back((uint32_t)(m >> 32));
                }
                p.task_rec.push_back((uint32_t)(ra | o << 8 | nodeLocal[c] << 16 | oc << 24));
            }
            while (p.inc.size() % kIncAlign) p.inc.push_back(0);
            incOff = (int)p.inc.size() - h.inc_start;
        }
        while (p.inc.size() % 8) p.inc.push_back(0);
        h.n_inc = (int)p.inc.size() - h.inc_start;
        h.n_task = (int)(p.task_rec.size() / 4) - h.task_start;
        p.tiles.push_back(h);
        p.max_node = std::max(p.max_node, h.n_node);
        p.max_elem = std::max(p.max_elem, TE);
        p.max_rows = std::max(p.max_rows, R);
        for (int i = h.node_start; i < (int)p.tile_nodes.size(); i++) nodeLocal[p.tile_nodes[i]] = -1;
        for (int e : tel) elemLocal[e] = -1;
        if (p.inc.size() > (size_t)INT32_MAX - 100000 || p.task_rec.size() > (size_t)INT32_MAX / 2)
            return MF_ERR_NOMEM;
    }
    return MF_OK;
}

int validate_plan(const PlanHost &p, const int *row, const int *col, const int *e2n,
                  int slot_bytes)
{
    const bool ela = p.op == MF_OP_ELA;
    if (p.tile_verts.size() != p.tile_elems.size() * 4) return MF_ERR_STATE;
    if (p.task_rec.size() % 4) return MF_ERR_STATE;
    const size_t nTasks = p.task_rec.size() / 4;
    std::vector<uint8_t> hit((size_t)p.nnz, 0), rowSeen((size_t)p.N, 0);
    for (int i : p.isolated) {
        if (i < 0 || i >= p.N || row[i + 1] != row[i] || rowSeen[i]) return MF_ERR_STATE;
        rowSeen[i] = 1;
    }
    std::vector<int> nodeStamp((size_t)p.N, -1), localOf((size_t)p.N, -1), cnt(kMaxTileNodes);
    std::vector<uint64_t> holds(kMaxTileNodes, 0);
    int tileNo = 0;
    for (const TileHdr &h : p.tiles) {
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
        tileNo++;
        for (int i = 0; i < h.n_node; i++) {
            const int n = nodes[i];
            if (n < 0 || n >= p.N || nodeStamp[n] == tileNo) return MF_ERR_STATE;
            nodeStamp[n] = tileNo;
            localOf[n] = i;
            if (i < R) {
                if (rowSeen[n]) return MF_ERR_STATE;
                rowSeen[n] = 1;
            }
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
            if (hit[(size_t)row[a] + doff]++) return MF_ERR_STATE;  // the diagonal block
            // the incidence list: every element of the tile holding a, ascending
            int last = -1;
            for (int i = 0; i < val; i++) {
                const uint16_t en = p.inc[(size_t)h.inc_start + incOff + i];
                const int t = en >> 2, va = en & 3;
                if (t >= h.n_elem || t <= last || verts[t * 4 + va] != ra) return MF_ERR_STATE;
                last = t;
                for (int k = 0; k < 4; k++) holds[verts[t * 4 + k]] |= 1ull << i;
            }
            if (cnt[ra] != val) return MF_ERR_STATE;
            // the row's tasks: exact masks, each block and its transpose written once
            for (; tk < tkEnd && (int)(p.task_rec[(size_t)tk * 4 + 3] & 0xFF) == ra; tk++) {
                const uint32_t *rec = p.task_rec.data() + (size_t)tk * 4;
                const int oa = (rec[3] >> 8) & 0xFF, cl = (rec[3] >> 16) & 0xFF, oc = rec[3] >> 24;
                if (oa >= len) return MF_ERR_STATE;
                const int c = col[row[a] + oa] - 1;
                if (c == a || hit[(size_t)row[a] + oa]++) return MF_ERR_STATE;
                const int ec = nodeStamp[c] == tileNo ? localOf[c] : -1;
                const int cnt = ec < 0 ? 0 : __builtin_popcountll(holds[ec]);
                if (ec < 0 || cl != ec || cnt < 1) return MF_ERR_STATE;
                if (cnt <= kListLen) {
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
                } else {
                    const uint64_t m = (uint64_t)rec[1] | (uint64_t)rec[2] << 32;
                    if (rec[0] != kMaskRecord || m != holds[ec]) return MF_ERR_STATE;
                }
                if (ec < R) {  // c is a row of this tile: the task also writes (c, a)
                    if (oc >= (p.tile_rowinfo[h.rinfo_start + ec].len_diag & 0xFFFF) ||
                        col[row[c] + oc] != a + 1 || hit[(size_t)row[c] + oc]++)
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
    }
    for (int i = 0; i < p.N; i++)
        if (!rowSeen[i]) return MF_ERR_STATE;
    for (int64_t q = 0; q < p.nnz; q++)
        if (hit[q] != 1) return MF_ERR_STATE;
    return MF_OK;
}

}  // namespace mf
