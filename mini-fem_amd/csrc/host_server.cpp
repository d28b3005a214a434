// The served form's request loop, native: resident meshes, an HBM arena the caller hands
// over, `slots` streams, admission by arrival time and memory, and one multi-plan launch
// (mf_plans_assemble) per group of requests that arrived together. Each request runs the
// iteration of src/FEM.cc:183-232 on its mesh with its own coefficient field or
// coordinates; the results stay in the arena until the request completes (and are copied
// to the caller's host buffers when it asks).
//
// Single-threaded: the loop polls the slots' completion events, frees what finished,
// then fills the free slots. Nothing here allocates device memory or synchronises the
// device while serving.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <map>
#include <new>
#include <thread>
#include <utility>
#include <vector>

#include "abi_guard.h"
#include "minifem.h"

namespace {

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First fit over [0, size): free blocks sorted by offset, neighbours merged on free.
class Arena {
public:
    explicit Arena(int64_t size) : size_(size) { free_.push_back({0, size}); }
    int64_t size() const { return size_; }
    int64_t used() const { return used_; }
    // offset of a block of n bytes (a multiple of kAlign), or -1
    int64_t alloc(int64_t n)
    {
        for (size_t i = 0; i < free_.size(); i++) {
            if (free_[i].second < n) continue;
            const int64_t off = free_[i].first;
            if (free_[i].second == n) free_.erase(free_.begin() + (int64_t)i);
            else free_[i] = {off + n, free_[i].second - n};
            used_ += n;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n)
    {
        auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(off, (int64_t)0));
        it = free_.insert(it, {off, n});
        auto nx = it + 1;
        if (nx != free_.end() && off + n == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = it - 1;
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        used_ -= n;
    }

private:
    int64_t size_, used_ = 0;
    std::vector<std::pair<int64_t, int64_t>> free_;
};

struct Admitted {
    int req;         // index into the run's requests
    int64_t off, n;  // arena block
};

}  // namespace

struct mf_server {
    unsigned char *arena_mem = nullptr;
    Arena arena;
    int max_group = 1, max_bypass = 0;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<mf_server_mesh> meshes;
    std::map<std::pair<int, std::vector<int>>, bool> fits;  // (mode, mesh ids) -> one launch

    mf_server(void *mem, int64_t bytes)
        : arena_mem(static_cast<unsigned char *>(mem)), arena(bytes)
    {
    }
    ~mf_server()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    }

    int64_t footprint(const mf_request &r) const
    {
        const mf_server_mesh &m = meshes[(size_t)r.mesh];
        return align_up(r.payload_bytes) + align_up(8 * m.nvalues) + align_up(8 * m.nprec);
    }

    mf_plan_job job(const mf_request &r, int64_t off) const
    {
        const mf_server_mesh &m = meshes[(size_t)r.mesh];
        unsigned char *inp = arena_mem + off;
        double *values = reinterpret_cast<double *>(inp + align_up(r.payload_bytes));
        double *prec = reinterpret_cast<double *>(reinterpret_cast<unsigned char *>(values) +
                                                  align_up(8 * m.nvalues));
        mf_plan_job j;
        j.plan = m.plan;
        j.coord = r.kind == MF_REQ_COORD ? reinterpret_cast<const double *>(inp) : m.coord;
        j.coef = r.kind == MF_REQ_COEF ? reinterpret_cast<const double *>(inp) : m.coef;
        j.values = values;
        j.prec = prec;
        j.singular = nullptr;
        return j;
    }

    // whether these meshes' plans run as one launch in this mode (cached per mesh set)
    bool group_fits(int mode, std::vector<int> ids)
    {
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        auto key = std::make_pair(mode, ids);
        auto it = fits.find(key);
        if (it != fits.end()) return it->second;
        std::vector<mf_plan_job> jobs;
        for (int i : ids) {
            const mf_server_mesh &m = meshes[(size_t)i];
            // the check reads only the plans; any non-null outputs will do
            jobs.push_back({m.plan, m.coord, m.coef, reinterpret_cast<double *>(arena_mem),
                            reinterpret_cast<double *>(arena_mem), nullptr});
        }
        const bool ok = mf_plans_lds_bytes(jobs.data(), (int)jobs.size(), mode) >= 0;
        fits.emplace(key, ok);
        return ok;
    }
};

extern "C" int mf_server_create(mf_server **server, void *arena, int64_t arena_bytes, int slots,
                                int max_group, int max_bypass)
{
    return mf_guarded([&]() -> int {
        if (!server || !arena || arena_bytes <= 0 || slots < 1 || slots > 16 || max_group < 1 ||
            max_group > MF_MAX_PLAN_JOBS || max_bypass < 0)
            return MF_ERR_ARG;
        *server = nullptr;
        auto *s = new mf_server(arena, arena_bytes);
        s->max_group = max_group;
        s->max_bypass = max_bypass;
        for (int i = 0; i < slots; i++) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr;
            if (hipStreamCreate(&st) != hipSuccess) {
                delete s;
                return MF_ERR_HIP;
            }
            s->streams.push_back(st);
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                delete s;
                return MF_ERR_HIP;
            }
            s->events.push_back(ev);
        }
        *server = s;
        return MF_OK;
    });
}

extern "C" int mf_server_add_mesh(mf_server *server, const mf_server_mesh *mesh)
{
    return mf_guarded([&]() -> int {
        if (!server || !mesh || !mesh->plan || !mesh->coord || mesh->nvalues <= 0 ||
            mesh->nprec <= 0 || mesh->ncoord <= 0 || mesh->ncoef <= 0)
            return MF_ERR_ARG;
        server->meshes.push_back(*mesh);
        return (int)server->meshes.size() - 1;
    });
}

extern "C" int64_t mf_server_arena_used(const mf_server *server)
{
    return server ? server->arena.used() : 0;
}

extern "C" int mf_server_destroy(mf_server *server)
{
    delete server;
    return MF_OK;
}

extern "C" int mf_server_run(mf_server *server, const mf_request *reqs, int n, double *t_start,
                             double *t_done, mf_server_stats *stats)
{
    return mf_guarded([&]() -> int {
        if (!server || n < 0 || (n > 0 && (!reqs || !t_start || !t_done))) return MF_ERR_ARG;
        mf_server &S = *server;
        const int nm = (int)S.meshes.size();
        for (int i = 0; i < n; i++) {
            const mf_request &r = reqs[i];
            if (r.mesh < 0 || r.mesh >= nm || (r.kind != MF_REQ_COEF && r.kind != MF_REQ_COORD) ||
                (r.mode != MF_PREC_FP64 && r.mode != MF_PREC_MIXED &&
                 r.mode != MF_PREC_MIXED_F32COEF) ||
                (r.mode == MF_PREC_MIXED_F32COEF && r.kind != MF_REQ_COEF) || !r.payload)
                return MF_ERR_ARG;
            const mf_server_mesh &m = S.meshes[(size_t)r.mesh];
            const int64_t want = r.kind == MF_REQ_COORD ? 8 * m.ncoord
                                 : r.mode == MF_PREC_MIXED_F32COEF ? 4 * m.ncoef : 8 * m.ncoef;
            if (r.payload_bytes != want) return MF_ERR_ARG;  // the kernel reads exactly this
            // the mesh's plan must launch in this mode (a plan staged in float serves the
            // mixed modes only), checked before anything runs
            if (!S.group_fits(r.mode, {r.mesh})) return MF_ERR_ARG;
            if (S.footprint(r) > S.arena.size()) return MF_ERR_BUDGET;
        }
        std::vector<int> pending((size_t)n);
        for (int i = 0; i < n; i++) pending[(size_t)i] = i;
        std::stable_sort(pending.begin(), pending.end(),
                         [&](int a, int b) { return reqs[a].t_arrival < reqs[b].t_arrival; });
        const int slots = (int)S.streams.size();
        std::vector<std::vector<Admitted>> busy((size_t)slots);
        std::vector<char> waited((size_t)n, 0);
        int64_t launches = 0, memory_waits = 0;
        int bypassed = 0, done_count = 0;
        size_t head = 0;  // pending[head..] not admitted yet, in arrival order, with holes
        std::vector<char> taken((size_t)n, 0);
        const auto t0 = std::chrono::steady_clock::now();
        auto now = [&]() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        std::vector<Admitted> group;
        // on an error: wait for what was issued, give every block back, report the status
        auto fail = [&](int status) {
            for (hipStream_t st : S.streams) (void)hipStreamSynchronize(st);
            for (auto &b : busy)
                for (const Admitted &a : b) S.arena.release(a.off, a.n);
            for (const Admitted &a : group) S.arena.release(a.off, a.n);
            return status;
        };
        double first_arrival = n > 0 ? reqs[pending[0]].t_arrival : 0.0, last_done = 0.0;
        std::vector<mf_plan_job> jobs;
        try {
            while (done_count < n) {
                bool progress = false;
                // 1: completions free their slot and arena blocks
                for (int k = 0; k < slots; k++) {
                    if (busy[(size_t)k].empty()) continue;
                    const hipError_t q = hipEventQuery(S.events[(size_t)k]);
                    if (q == hipErrorNotReady) continue;
                    if (q != hipSuccess) return fail(MF_ERR_HIP);
                    const double t = now();
                    for (const Admitted &a : busy[(size_t)k]) {
                        t_done[a.req] = t;
                        S.arena.release(a.off, a.n);
                        done_count++;
                    }
                    last_done = t;
                    busy[(size_t)k].clear();
                    progress = true;
                }
                // 2: free slots take the requests that have arrived, in arrival order: one
                // precision mode per launch, plans that fit one launch, memory first fit
                const double tn = now();
                while (head < pending.size() && taken[(size_t)pending[head]]) head++;
                for (int k = 0; k < slots; k++) {
                    if (!busy[(size_t)k].empty()) continue;
                    if (head >= pending.size() || reqs[pending[head]].t_arrival > tn) break;
                    group.clear();
                    std::vector<int> ids;
                    int mode = -1;
                    const int hreq = pending[head];
                    bool head_taken = false;
                    for (size_t p = head; p < pending.size() && (int)group.size() < S.max_group;
                         p++) {
                        const int ri = pending[p];
                        if (taken[(size_t)ri]) continue;
                        const mf_request &r = reqs[ri];
                        if (r.t_arrival > tn) break;
                        if (mode >= 0) {
                            if (r.mode != mode) continue;
                            ids.push_back(r.mesh);
                            const bool ok = S.group_fits(mode, ids);
                            ids.pop_back();
                            if (!ok) continue;
                        }
                        const int64_t need = S.footprint(r);
                        const int64_t off = S.arena.alloc(need);
                        if (off < 0) {
                            if (!waited[(size_t)ri]) {
                                waited[(size_t)ri] = 1;
                                memory_waits++;
                            }
                            // the head has waited long enough: nothing passes it
                            if (ri == hreq && bypassed >= S.max_bypass) break;
                            continue;
                        }
                        taken[(size_t)ri] = 1;
                        if (ri == hreq) head_taken = true;
                        group.push_back({ri, off, need});
                        ids.push_back(r.mesh);
                        mode = r.mode;
                    }
                    // nothing admissible until memory frees up or more requests arrive
                    if (group.empty()) break;
                    if (head_taken) bypassed = 0;
                    else bypassed++;  // the head stays behind the requests just admitted
                    hipStream_t st = S.streams[(size_t)k];
                    jobs.clear();
                    for (const Admitted &a : group) {
                        const mf_request &r = reqs[a.req];
                        t_start[a.req] = tn;
                        if (hipMemcpyAsync(S.arena_mem + a.off, r.payload, (size_t)r.payload_bytes,
                                           hipMemcpyHostToDevice, st) != hipSuccess)
                            return fail(MF_ERR_HIP);
                        jobs.push_back(S.job(r, a.off));
                    }
                    const int stl = mf_plans_assemble(jobs.data(), (int)jobs.size(), mode, st);
                    if (stl != MF_OK) return fail(stl);
                    for (size_t g = 0; g < group.size(); g++) {
                        const mf_request &r = reqs[group[g].req];
                        const mf_server_mesh &m = S.meshes[(size_t)r.mesh];
                        if ((r.values_out &&
                             hipMemcpyAsync(r.values_out, jobs[g].values, (size_t)(8 * m.nvalues),
                                            hipMemcpyDeviceToHost, st) != hipSuccess) ||
                            (r.prec_out &&
                             hipMemcpyAsync(r.prec_out, jobs[g].prec, (size_t)(8 * m.nprec),
                                            hipMemcpyDeviceToHost, st) != hipSuccess))
                            return fail(MF_ERR_HIP);
                    }
                    if (hipEventRecord(S.events[(size_t)k], st) != hipSuccess)
                        return fail(MF_ERR_HIP);
                    busy[(size_t)k].swap(group);
                    group.clear();
                    launches++;
                    progress = true;
                    while (head < pending.size() && taken[(size_t)pending[head]]) head++;
                }
                if (!progress) std::this_thread::yield();
            }
        } catch (const std::bad_alloc &) {  // a host container: nothing may stay in flight
            return fail(MF_ERR_NOMEM);
        } catch (...) {
            return fail(MF_ERR_STATE);
        }
        if (stats) {
            stats->launches = launches;
            stats->memory_waits = memory_waits;
            stats->makespan_s = n > 0 ? last_done - first_arrival : 0.0;
        }
        return MF_OK;
    });
}
