// Host construction and validation of the tiled gather schedule (see plan.h).
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
    int N, nE, max_rows, budget, eb, bb;
    const int *e2n, *row;
    const std::vector<int> &n2eI, &n2eV;
    const double *coord;
    std::vector<int> order;
    std::vector<std::pair<int, int>> leaves;
    std::vector<int> elemStamp, nodeStamp;
    int stampId = 0;

    Partitioner(int N_, int nE_, int mr, int bud, int eb_, int bb_, const int *e2n_,
                const int *row_, const std::vector<int> &I, const std::vector<int> &V,
                const double *c)
        : N(N_), nE(nE_), max_rows(mr), budget(bud), eb(eb_), bb(bb_), e2n(e2n_), row(row_),
          n2eI(I), n2eV(V), coord(c), order(N_), elemStamp(nE_, -1), nodeStamp(N_, -1)
    {
        std::iota(order.begin(), order.end(), 0);
    }

    // Exact LDS bytes of a tile made of rows order[lo, hi), or INT32_MAX when it exceeds
    // the element, node or row limits.
    int64_t lds_of(int lo, int hi)
    {
        stampId++;
        for (int i = lo; i < hi; i++) nodeStamp[order[i]] = stampId;
        int64_t TE = 0, NN = hi - lo;
        for (int i = lo; i < hi; i++) {
            const int r = order[i];
            if (row[r + 1] - row[r] > kMaxRowLen) return INT32_MAX;
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
        if (NN > kMaxTileNodes || TE > kMaxTileElems || hi - lo > kMaxTileRows) return INT32_MAX;
        return tile_lds_layout((int)NN, (int)TE, hi - lo, eb, bb).total;
    }

    bool fits(int lo, int hi) { return hi - lo <= max_rows && lds_of(lo, hi) <= budget; }

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
        const int mid = lo + (hi - lo) / 2;
        const double *c = coord;
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi,
                         [c, ax](int a, int b) {
                             const double xa = c[(int64_t)a * 3 + ax], xb = c[(int64_t)b * 3 + ax];
                             return xa < xb || (xa == xb && a < b);
                         });
        rcb(lo, mid);
        rcb(mid, hi);
    }

    // Breadth-first order of the node graph (through shared elements), then greedy
    // chunks that fit.
    void bfs_chunks()
    {
        std::vector<char> seen(N, 0);
        int head = 0, tail = 0;
        for (int s = 0; s < N; s++) {
            if (seen[s]) continue;
            seen[s] = 1;
            order[tail++] = s;
            while (head < tail) {
                const int r = order[head++];
                for (int j = n2eI[r]; j < n2eI[r + 1]; j++) {
                    const int e = n2eV[j];
                    for (int k = 0; k < 4; k++) {
                        const int v = e2n[(int64_t)e * 4 + k] - 1;
                        if (!seen[v]) { seen[v] = 1; order[tail++] = v; }
                    }
                }
            }
        }
        int lo = 0;
        while (lo < N) {
            int hi = lo + 1;
            while (hi < N && fits(lo, hi + 1)) hi++;
            leaves.emplace_back(lo, hi);
            lo = hi;
        }
    }
};

// Appends one chunk's contributions in the grouped, lane-interleaved layout.
void emit_chunk(PlanHost &p, int group, int te, const std::vector<const uint16_t *> &lists,
                const std::vector<int> &counts)
{
    ChunkHdr c{};
    c.contrib_start = (int)p.contribs.size();
    c.steps = 0;
    for (int n : counts) c.steps = std::max(c.steps, n);
    const int groups = (c.steps + group - 1) / group;
    for (int g = 0; g < groups; g++)
        for (int l = 0; l < kTaskLanes; l++)
            for (int j = 0; j < group; j++) {
                const int s = g * group + j;
                uint16_t en = (uint16_t)(te << 4);  // padding: the zero element slot
                if (l < (int)counts.size() && s < counts[l]) en = lists[l][s];
                p.contribs.push_back(en);
            }
    p.chunks.push_back(c);
}

}  // namespace

int build_plan(PlanHost &p, int op, int nE, int N, int nnz, const int *e2n, const int *row,
               const int *col, const int *checkBounds, const double *coord, int lds_budget,
               int max_rows, int slot_bytes)
{
    p = PlanHost{};
    p.op = op;
    p.nE = nE;
    p.N = N;
    p.nnz = nnz;
    if (nE < 0 || N < 0 || nnz < 0 || max_rows < 1 || max_rows > kMaxTileRows) return MF_ERR_ARG;
    if (row[0] != 0 || row[N] != nnz) return MF_ERR_ARG;
    for (int i = 0; i < N; i++) {
        const int len = row[i + 1] - row[i];
        if (len < 0 || len > kMaxRowLen) return MF_ERR_ARG;
        if (len == 0) p.isolated.push_back(i);
    }
    std::vector<int> n2eI((size_t)N + 1, 0), n2eV((size_t)nE * 4 + 1);
    for (int64_t q = 0; q < (int64_t)nE * 4; q++) {
        const int v = e2n[q] - 1;
        if (v < 0 || v >= N) return MF_ERR_ARG;
        n2eI[v + 1]++;
    }
    for (int i = 0; i < N; i++) n2eI[i + 1] += n2eI[i];
    {
        std::vector<int> cur(n2eI.begin(), n2eI.end() - 1);
        for (int e = 0; e < nE; e++)
            for (int k = 0; k < 4; k++) n2eV[cur[e2n[(int64_t)e * 4 + k] - 1]++] = e;
    }
    const int bb = op == MF_OP_ELA ? 72 : 8;
    Partitioner part(N, nE, max_rows, lds_budget, slot_bytes, bb, e2n, row, n2eI, n2eV, coord);
    if (coord) part.rcb(0, N);
    else part.bfs_chunks();

    std::vector<int> nodeLocal((size_t)N, -1), elemLocal((size_t)nE, -1);
    std::vector<int> rows, tel;
    // per-row contribution lists of the tile: diag list + one list per CSR entry
    std::vector<uint16_t> buf;
    std::vector<int> diagStart, diagCnt, entStart, entCnt;
    std::vector<int> dorder;
    std::vector<const uint16_t *> lists;
    std::vector<int> counts;

    for (const auto &lf : part.leaves) {
        rows.assign(part.order.begin() + lf.first, part.order.begin() + lf.second);
        std::sort(rows.begin(), rows.end());
        const int R = (int)rows.size();
        if (R > kMaxTileRows) return MF_ERR_ARG;
        TileHdr h{};
        h.node_start = (int)p.tile_nodes.size();
        h.n_rows = R;
        h.rinfo_start = (int)p.tile_rowinfo.size();
        for (int i = 0; i < R; i++) {
            const int a = rows[i];
            nodeLocal[a] = i;
            p.tile_nodes.push_back(a);
            int mask = 0;
            if (op == MF_OP_ELA && checkBounds)
                for (int k = 0; k < 3; k++)
                    if (checkBounds[(int64_t)k * N + a] != 0) mask |= 1 << k;
            p.tile_rowinfo.push_back(RowInfo{row[a], (row[a + 1] - row[a]) | mask << 16});
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

        // contribution lists, ascending tile element order (= ascending element id)
        buf.clear();
        diagStart.assign(R, 0);
        diagCnt.assign(R, 0);
        entStart.clear();
        entCnt.clear();
        for (int ra = 0; ra < R; ra++) {
            const int a = rows[ra];
            diagStart[ra] = (int)buf.size();
            for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                const int e = n2eV[j];
                const int *v = e2n + (int64_t)e * 4;
                int va = 0;
                for (int k = 0; k < 4; k++)
                    if (v[k] - 1 == a) va = k;
                buf.push_back((uint16_t)(elemLocal[e] << 4 | va << 2 | va));
            }
            diagCnt[ra] = (int)buf.size() - diagStart[ra];
            for (int pa = row[a]; pa < row[a + 1]; pa++) {
                const int c = col[pa] - 1;
                entStart.push_back((int)buf.size());
                if (c != a) {
                    for (int j = n2eI[a]; j < n2eI[a + 1]; j++) {
                        const int e = n2eV[j];
                        const int *v = e2n + (int64_t)e * 4;
                        int va = -1, vc = -1;
                        for (int k = 0; k < 4; k++) {
                            if (v[k] - 1 == a) va = k;
                            if (v[k] - 1 == c) vc = k;
                        }
                        if (vc < 0) continue;
                        buf.push_back((uint16_t)(elemLocal[e] << 4 | va << 2 | vc));
                    }
                }
                entCnt.push_back((int)buf.size() - entStart.back());
            }
        }

        // B1 diagonal chunks: rows with a diagonal, by contribution count (balanced waves)
        dorder.clear();
        for (int ra = 0; ra < R; ra++)
            if (row[rows[ra] + 1] > row[rows[ra]]) dorder.push_back(ra);
        std::stable_sort(dorder.begin(), dorder.end(),
                         [&](int x, int y) { return diagCnt[x] > diagCnt[y]; });
        h.dchunk_start = (int)p.chunks.size();
        for (size_t i = 0; i < dorder.size(); i += kTaskLanes) {
            lists.clear();
            counts.clear();
            for (size_t l = 0; l < (size_t)kTaskLanes; l++) {
                if (i + l < dorder.size()) {
                    const int ra = dorder[i + l];
                    lists.push_back(buf.data() + diagStart[ra]);
                    counts.push_back(diagCnt[ra]);
                    p.tasks.push_back((uint32_t)ra);
                } else {
                    p.tasks.push_back(kNoTask);
                }
            }
            emit_chunk(p, kDiagGroup, TE, lists, counts);
        }
        h.n_dchunks = (int)p.chunks.size() - h.dchunk_start;

        // B2 row chunks: whole rows in tile-row order, one lane per CSR entry; a row longer
        // than the lanes left starts a new chunk (or spans chunks if longer than 64)
        h.rchunk_start = (int)p.chunks.size();
        lists.clear();
        counts.clear();
        std::vector<uint32_t> words;
        auto flush = [&]() {
            if (words.empty()) return;
            for (int l = 0; l < kTaskLanes; l++)
                p.tasks.push_back(l < (int)words.size() ? words[l] : kNoTask);
            emit_chunk(p, kRowGroup, TE, lists, counts);
            lists.clear();
            counts.clear();
            words.clear();
        };
        int ent = 0;
        for (int ra = 0; ra < R; ra++) {
            const int a = rows[ra], len = row[a + 1] - row[a];
            if (len <= kTaskLanes && (int)words.size() + len > kTaskLanes) flush();
            for (int oa = 0; oa < len; oa++, ent++) {
                if ((int)words.size() == kTaskLanes) flush();
                const bool diag = col[row[a] + oa] - 1 == a;
                words.push_back((uint32_t)ra | (uint32_t)oa << 8 | (diag ? kDiagLane : 0u));
                lists.push_back(buf.data() + entStart[ent]);
                counts.push_back(entCnt[ent]);
            }
        }
        flush();
        h.n_rchunks = (int)p.chunks.size() - h.rchunk_start;

        p.total_tasks += (int64_t)dorder.size() + ent;
        p.tiles.push_back(h);
        p.max_node = std::max(p.max_node, h.n_node);
        p.max_elem = std::max(p.max_elem, TE);
        p.max_rows = std::max(p.max_rows, R);
        for (int i = h.node_start; i < (int)p.tile_nodes.size(); i++) nodeLocal[p.tile_nodes[i]] = -1;
        for (int e : tel) elemLocal[e] = -1;
        if (p.contribs.size() > (size_t)INT32_MAX - 1000000) return MF_ERR_NOMEM;
    }
    return MF_OK;
}

int validate_plan(const PlanHost &p, const int *row, const int *col, int slot_bytes)
{
    std::vector<uint8_t> hit((size_t)p.nnz, 0);
    std::vector<uint8_t> rowSeen((size_t)p.N, 0);
    const int bb = p.op == MF_OP_ELA ? 72 : 8;
    if (p.tasks.size() != p.chunks.size() * kTaskLanes) return MF_ERR_STATE;
    if (p.tile_verts.size() != p.tile_elems.size() * 4) return MF_ERR_STATE;
    for (const TileHdr &h : p.tiles) {
        if (h.n_rows < 1 || h.n_rows > kMaxTileRows || h.n_node < h.n_rows ||
            h.n_node > kMaxTileNodes || h.n_elem < 0 || h.n_elem > kMaxTileElems)
            return MF_ERR_STATE;
        if (h.node_start < 0 || (size_t)h.node_start + h.n_node > p.tile_nodes.size()) return MF_ERR_STATE;
        if (h.elem_start < 0 || (size_t)h.elem_start + h.n_elem > p.tile_elems.size()) return MF_ERR_STATE;
        if (h.dchunk_start < 0 || h.n_dchunks < 0 || h.rchunk_start < 0 || h.n_rchunks < 0 ||
            (size_t)h.dchunk_start + h.n_dchunks > p.chunks.size() ||
            (size_t)h.rchunk_start + h.n_rchunks > p.chunks.size())
            return MF_ERR_STATE;
        if (tile_lds_layout(h.n_node, h.n_elem, h.n_rows, slot_bytes, bb).total + 16 > kMaxLdsBytes)
            return MF_ERR_STATE;
        for (int i = 0; i < h.n_node; i++) {
            const int n = p.tile_nodes[h.node_start + i];
            if (n < 0 || n >= p.N) return MF_ERR_STATE;
            if (i < h.n_rows) {
                if (rowSeen[n]) return MF_ERR_STATE;
                rowSeen[n] = 1;
            }
        }
        for (int t = 0; t < h.n_elem; t++) {
            const int e = p.tile_elems[h.elem_start + t];
            if (e < 0 || e >= p.nE) return MF_ERR_STATE;
            if (t > 0 && e <= p.tile_elems[h.elem_start + t - 1]) return MF_ERR_STATE;
            for (int k = 0; k < 4; k++)
                if (p.tile_verts[(size_t)(h.elem_start + t) * 4 + k] >= h.n_node) return MF_ERR_STATE;
        }
        if (h.rinfo_start < 0 || (size_t)h.rinfo_start + h.n_rows > p.tile_rowinfo.size())
            return MF_ERR_STATE;
        for (int i = 0; i < h.n_rows; i++) {
            const int n = p.tile_nodes[h.node_start + i];
            const RowInfo ri = p.tile_rowinfo[h.rinfo_start + i];
            if (ri.rowptr != row[n] || (ri.len_mask & 0xFFFF) != row[n + 1] - row[n] ||
                (ri.len_mask >> 16) > 7)
                return MF_ERR_STATE;
        }
        std::vector<int> diagSeen(h.n_rows, 0);
        for (int pass = 0; pass < 2; pass++) {
            const int c0 = pass == 0 ? h.dchunk_start : h.rchunk_start;
            const int nc = pass == 0 ? h.n_dchunks : h.n_rchunks;
            const int G = pass == 0 ? kDiagGroup : kRowGroup;
            for (int ch = c0; ch < c0 + nc; ch++) {
                const ChunkHdr &c = p.chunks[ch];
                const int groups = (c.steps + G - 1) / G;
                if (c.steps < 0 || c.contrib_start < 0 || c.contrib_start % G ||
                    (int64_t)c.contrib_start + (int64_t)groups * G * kTaskLanes >
                        (int64_t)p.contribs.size())
                    return MF_ERR_STATE;
                for (int l = 0; l < kTaskLanes; l++) {
                    int last = -1;
                    for (int s = 0; s < groups * G; s++) {
                        const uint16_t en = p.contribs[(size_t)c.contrib_start +
                                                       (size_t)(s / G) * G * kTaskLanes +
                                                       l * G + s % G];
                        const int t = en >> 4;
                        if (t == h.n_elem) {  // padding, only after the lane's real entries
                            if (en != (uint16_t)(h.n_elem << 4)) return MF_ERR_STATE;
                            last = INT_MAX;
                            continue;
                        }
                        if (t > h.n_elem || t <= last) return MF_ERR_STATE;
                        last = t;
                    }
                    const uint32_t w = p.tasks[(size_t)ch * kTaskLanes + l];
                    if (w == kNoTask) continue;
                    if (pass == 0) {
                        if ((int)w >= h.n_rows || diagSeen[w]++) return MF_ERR_STATE;
                        const int a = p.tile_nodes[h.node_start + w];
                        if (row[a + 1] == row[a]) return MF_ERR_STATE;
                        continue;
                    }
                    const int ra = w & 0xFF, oa = (w >> 8) & 0xFF;
                    if ((w & ~(kDiagLane | 0xFFFFu)) || ra >= h.n_rows) return MF_ERR_STATE;
                    const int a = p.tile_nodes[h.node_start + ra];
                    if (oa >= row[a + 1] - row[a]) return MF_ERR_STATE;
                    const int pos = row[a] + oa;
                    if (hit[pos]++) return MF_ERR_STATE;
                    if (((w & kDiagLane) != 0) != (col[pos] == a + 1)) return MF_ERR_STATE;
                }
            }
        }
        for (int ra = 0; ra < h.n_rows; ra++) {
            const int a = p.tile_nodes[h.node_start + ra];
            if ((row[a + 1] > row[a]) != (diagSeen[ra] == 1)) return MF_ERR_STATE;
        }
    }
    for (int i = 0; i < p.N; i++)
        if (!rowSeen[i]) return MF_ERR_STATE;
    for (int64_t q = 0; q < p.nnz; q++)
        if (hit[q] != 1) return MF_ERR_STATE;
    return MF_OK;
}

}  // namespace mf
