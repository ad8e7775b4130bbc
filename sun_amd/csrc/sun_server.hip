// sun_server.hip -- continuous-batching server for ensemble-generation jobs.
//
// Admission: a job is admitted when its link field and scratch (lattice_footprint)
// fit in the free part of the HBM budget, which is reserved once at creation and
// sub-allocated by an Arena, so no hipMalloc/hipFree happens while jobs run.
// Pending jobs are taken in arrival order; a job that does not fit is passed over
// for later, smaller ones ("backfill") unless it has already been passed over
// kMaxPassOver times, in which case admission stops until it fits.
//
// Execution: resident jobs are grouped by kind (kernel variant, n_or, reunit) and,
// for large lattices, by volume.  A group runs its iterations with the batched
// kernels -- one launch per (parity, mu) for all its lattices; for large lattices the
// single-lattice kernels, one launch per member -- captured once per membership into
// a hipGraph, on its own stream.  Groups progress independently:
// the scheduling thread launches a few iterations of an idle group (as many as take
// about kTargetMs, never more than any member still needs), polls the groups'
// completion events, and at each completion retires finished jobs (history copied
// out, memory freed), admits pending ones and re-captures the group.  A job's sweep
// counter and observable history live in its batch-table entries and are advanced
// by the kernels themselves.
//
// A job's arithmetic depends only on its own lattice, seed and sweep number (the
// per-site update and the per-lattice fixed-order reductions are the same code
// alone and batched), so its results are bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "sun_arena.hpp"
#include "sun_internal.hpp"

using namespace sun;

namespace {

constexpr int kMaxPassOver = 8;
// work per group launch before the scheduler looks again.  Between two launches of a
// group the device idles for the completion poll, the retire/admit step and the graph
// launch (~0.1 ms): 7% of a served 32^4 job's time at 2 ms per launch, 1% at 10 ms,
// 0.25% at 40 ms.  A job arriving for a running group joins at its next launch.
constexpr double kTargetMs = 40.0;
constexpr int kMaxItersPerLaunch = 64;
constexpr int kLargeVolume = 1 << 17;  // lattices at least this large get a group per volume

struct Group;

struct Job {
    uint64_t id = 0;
    sun_job_desc d{};
    int state = SUN_JOB_PENDING;
    int iters_done = 0;
    int passed_over = 0;
    bool cancel = false;
    double t_submit = 0, t_start = 0, t_end = 0;
    size_t off = SIZE_MAX, bytes = 0;
    sun_lattice* lat = nullptr;
    Group* group = nullptr;
    std::vector<double2> hist;  // host copy of the history once done
    // suspension requested (running jobs: honoured at the next iteration boundary)
    bool suspend = false;
    std::string suspend_path;
    // a resumed job: field bytes and the observable rows of its first iters_done
    // iterations, copied to its memory at admission (kept until it retires)
    std::vector<unsigned char> resume_field;
    std::vector<double2> resume_hist;
    std::string error;  // why the job failed (state SUN_JOB_FAILED)
};

bool finished(const Job* J) {
    return J->state == SUN_JOB_DONE || J->state == SUN_JOB_CANCELLED || J->state == SUN_JOB_SUSPENDED ||
           J->state == SUN_JOB_FAILED;
}

struct Group {
    const OpsTable* ops = nullptr;
    int n_or = 0, reunit = 0, vkey = 0;
    int piped[2] = {0, 0};  // staple kernel of the heatbath / overrelaxation launches
    int brick[2] = {0, 0};  // their members' update bricks instead of linear order
    int lds[2] = {0, 0};    // unused dynamic LDS per workgroup (occupancy cap)
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int cap = 0;
    BatchEntry *t_update = nullptr, *t_reunit = nullptr, *t_plaq = nullptr, *t_poly = nullptr;
    double2 *d_plaq = nullptr, *d_poly = nullptr;
    std::vector<Job*> members;   // in the device tables
    std::vector<Job*> incoming;  // admitted (start kernels queued), join at the next launch
    bool dirty = true;           // tables and graph need rebuilding
    int blocks_update = 0, blocks_reunit = 0, blocks_plaq = 0, blocks_poly = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int inflight = 0;            // iterations launched and not yet completed
    double t_launch = 0, iter_ms = 0;
    // Large-volume groups (every member has the same volume >= kLargeVolume): each
    // member's update launches go to the single-lattice kernels, which take the field
    // and geometry as kernel arguments instead of reading a table entry (measured at
    // 32^4: 9% faster launches, each still fills the GPU); the heatbath sweep counter
    // is read from the member's table entry.  Same code per site, so the links are
    // bit-identical to the batched kernels'.
    struct Solo {
        Field f;
        Geom g;
        uint64_t seed = 0;
        double bnc = 0;
        const uint32_t* sweep_dev = nullptr;
        double2* partials = nullptr;
        LaunchCfg cfg[2];
    };
    std::vector<Solo> solo;
    // update chunks (kCacheChunkBytes): consecutive members whose link
    // fields together fit the memory-side cache run their heatbath and overrelaxation
    // launches back to back; t_update's first_block counts from the chunk's start
    struct Chunk {
        int first = 0, n = 0, blocks = 0;
    };
    std::vector<Chunk> chunks;
};

double now_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct sun_server {
    sun_ctx* ctx = nullptr;
    size_t budget = 0;
    int max_jobs = 0;
    char* base = nullptr;
    sun::Arena arena;
    std::mutex mu;  // jobs, pending, running, arena, counters, job states
    std::condition_variable cv;
    std::deque<Job*> pending;
    std::map<uint64_t, Job*> jobs;
    std::vector<Job*> running;
    // groups are driven by one thread at a time: the scheduling thread, or
    // sun_server_step's caller while that thread is off (step_mu)
    std::vector<Group*> groups;
    std::thread worker;
    bool stop = false, worker_on = false, admit_needed = false;
    int worker_rc = SUN_OK;
    uint64_t next_id = 1;
    std::chrono::steady_clock::time_point t0;
    uint64_t jobs_done = 0, link_updates = 0, steps = 0;
    double busy_s = 0;
    std::mutex step_mu;
};

namespace {

void drop_graph(Group* g) {
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
}

void free_group(Group* g) {
    drop_graph(g);
    if (g->done) hipEventDestroy(g->done);
    if (g->stream) hipStreamDestroy(g->stream);
    hipFree(g->t_update);
    hipFree(g->t_reunit);
    hipFree(g->t_plaq);
    hipFree(g->t_poly);
    hipFree(g->d_plaq);
    hipFree(g->d_poly);
    delete g;
}

int group_vkey(const sun_lattice* L) { return L->g.V >= kLargeVolume ? L->g.V : 0; }

int make_group(sun_server* S, const sun_lattice* L, int n_or, int reunit, Group** out) {
    Group* g = new (std::nothrow) Group();
    if (!g) return fail(SUN_ERR_NOMEM, "server: host allocation");
    g->ops = L->ops;
    for (int k = 0; k < 2; k++) {
        g->piped[k] = default_batch_piped(L->ops, k);
        g->brick[k] = default_batch_brick(L->ops, k);
        g->lds[k] = default_batch_lds(L->ops, k);
    }
    g->n_or = n_or;
    g->reunit = reunit;
    g->vkey = group_vkey(L);
    g->cap = S->max_jobs;
    const size_t tb = sizeof(BatchEntry) * (size_t)g->cap;
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&g->t_update, tb);
    if (e == hipSuccess) e = hipMalloc(&g->t_reunit, tb);
    if (e == hipSuccess) e = hipMalloc(&g->t_plaq, tb);
    if (e == hipSuccess) e = hipMalloc(&g->t_poly, tb);
    if (e == hipSuccess) e = hipMalloc(&g->d_plaq, sizeof(double2) * (size_t)g->cap);
    if (e == hipSuccess) e = hipMalloc(&g->d_poly, sizeof(double2) * (size_t)g->cap);
    if (e != hipSuccess) {
        hipGetLastError();
        free_group(g);
        return fail(SUN_ERR_NOMEM, "server: group tables: %s", hipGetErrorString(e));
    }
    *out = g;
    return SUN_OK;
}

// device tables of an idle group from its members' host state
int rebuild_tables(Group* g) {
    const int n = (int)g->members.size();
    std::vector<BatchEntry> tu(n), tr(n), tp(n), tl(n);
    int cu = 0, cr = 0, cp = 0, cl = 0;
    size_t chunk_bytes = 0;
    g->chunks.clear();
    g->blocks_update = 0;
    for (int i = 0; i < n; i++) {
        Job* J = g->members[i];
        sun_lattice* L = J->lat;
        if (g->chunks.empty() || (g->chunks.back().n > 0 && chunk_bytes + L->bytes > kCacheChunkBytes)) {
            g->chunks.push_back(Group::Chunk{i, 0, 0});
            chunk_bytes = 0;
            cu = 0;
        }
        chunk_bytes += L->bytes;
        BatchEntry e;
        memset(&e, 0, sizeof(e));
        e.f = L->f;
        e.g = L->g;
        e.seed = J->d.seed;
        e.bnc = L->d.beta / L->d.nc;
        e.sweep = (uint32_t)J->iters_done;
        e.hist = L->hist;
        e.hist_iter = (uint32_t)J->iters_done;
        batch_brick(L->g, e.ub1, e.ub2, e.ub3, e.spt);
        BatchEntry a = e, r = e, p = e, q = e;
        a.first_block = cu;
        a.nblocks = update_blocks(L->g, 256);
        cu += a.nblocks;
        g->chunks.back().n++;
        g->chunks.back().blocks = cu;
        g->blocks_update += a.nblocks;
        r.first_block = cr;
        r.nblocks = link_blocks(L->g);
        cr += r.nblocks;
        p.first_block = cp;
        p.nblocks = plaquette_blocks(L->g);
        p.partials = L->partials;
        p.scale = 1.0 / ((double)(3 * L->d.nc) * (double)L->g.V);  // plaquette_cub.cu:234
        cp += p.nblocks;
        q.first_block = cl;
        q.nblocks = polyakov_blocks(L->g);
        // reused: the Polyakov launch follows the plaquette finalize; large-volume groups
        // measure both in one launch, Polyakov partials after the plaquette's
        q.partials = g->vkey != 0 ? L->partials + plaquette_blocks(L->g) : L->partials;
        q.scale = 1.0 / ((double)L->d.nc * (double)L->g.X[0] * L->g.X[1] * L->g.X[2]);  // polyakovloop.cu:651
        cl += q.nblocks;
        tu[i] = a;
        tr[i] = r;
        tp[i] = p;
        tl[i] = q;
    }
    g->solo.clear();
    if (g->vkey != 0) {
        for (int i = 0; i < n; i++) {
            const BatchEntry& a = tu[i];
            Group::Solo m;
            m.f = a.f;
            m.g = a.g;
            m.seed = a.seed;
            m.bnc = a.bnc;
            m.partials = tp[i].partials;
            m.sweep_dev = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(g->t_update + i) +
                                                            offsetof(BatchEntry, sweep));
            const sun_lattice* L = g->members[i]->lat;
            for (int k = 0; k < 2; k++) {
                // the context's shape for this lattice description when it has one
                // (autotuner cache or sun_ctx_set_launch), else the batch defaults
                const char* name = k == 0 ? "heatbath" : "overrelax";
                if (L->ctx->tuner.has(name, L->d)) {
                    m.cfg[k] = L->ctx->tuner.lookup(name, L->d);
                    continue;
                }
                LaunchCfg c;
                c.block = 256;
                c.lds = g->lds[k];
                c.piped = g->piped[k];
                if (g->brick[k] && a.ub1 >= 0 && a.spt == 1) {  // single-lattice launches: 256-thread bricks
                    c.b1 = a.ub1;
                    c.b2 = a.ub2;
                    c.b3 = a.ub3;
                }
                m.cfg[k] = c;
            }
            g->solo.push_back(m);
        }
    }
    g->blocks_reunit = cr;
    g->blocks_plaq = cp;
    g->blocks_poly = cl;
    const size_t b = sizeof(BatchEntry) * (size_t)n;
    if (n) {
        HIP_TRY(hipMemcpyAsync(g->t_update, tu.data(), b, hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipMemcpyAsync(g->t_reunit, tr.data(), b, hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipMemcpyAsync(g->t_plaq, tp.data(), b, hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipMemcpyAsync(g->t_poly, tl.data(), b, hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));  // the host vectors go out of scope
    }
    return SUN_OK;
}

// one iteration of every member of a group, enqueued on the group's stream
int group_iteration(Group* g) {
    const int n = (int)g->members.size();
    hipStream_t s = g->stream;
    // chunk by chunk (members are independent: each sees its launches in the same order)
    for (const Group::Chunk& c : g->chunks) {
        if (!g->solo.empty()) {
            for (int parity = 0; parity < 2; parity++)
                for (int mu = 0; mu < 4; mu++)
                    for (int i = c.first; i < c.first + c.n; i++) {
                        const Group::Solo& m = g->solo[i];
                        HIP_TRY(g->ops->update(m.f, m.g, parity, mu, true, 0, m.sweep_dev, m.bnc, m.seed, m.cfg[0], s));
                    }
            for (int k = 0; k < g->n_or; k++)
                for (int parity = 0; parity < 2; parity++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int i = c.first; i < c.first + c.n; i++) {
                            const Group::Solo& m = g->solo[i];
                            HIP_TRY(g->ops->update(m.f, m.g, parity, mu, false, 0, nullptr, 0.0, 0, m.cfg[1], s));
                        }
        } else {
            for (int parity = 0; parity < 2; parity++)
                for (int mu = 0; mu < 4; mu++)
                    HIP_TRY(g->ops->batch_update(g->t_update + c.first, c.n, c.blocks, parity, mu, true, g->piped[0],
                                                 g->brick[0], g->lds[0], s));
            for (int k = 0; k < g->n_or; k++)
                for (int parity = 0; parity < 2; parity++)
                    for (int mu = 0; mu < 4; mu++)
                        HIP_TRY(g->ops->batch_update(g->t_update + c.first, c.n, c.blocks, parity, mu, false,
                                                     g->piped[1], g->brick[1], g->lds[1], s));
        }
    }
    if (g->reunit) HIP_TRY(g->ops->batch_reunitarize(g->t_reunit, n, g->blocks_reunit, s));
    if (!g->solo.empty()) {
        // plaquette and Polyakov partials of a member in one launch (the same partial
        // sums as the two batched kernels)
        for (const Group::Solo& m : g->solo) HIP_TRY(g->ops->measure_partial(m.f, m.g, m.partials, s));
        // both finalizes, the history record and the sweep advance in one launch
        HIP_TRY(launch_group_close(g->t_plaq, g->t_poly, g->t_update, n, g->d_plaq, g->d_poly, s));
        return SUN_OK;
    }
    HIP_TRY(g->ops->batch_plaquette_partial(g->t_plaq, n, g->blocks_plaq, s));
    HIP_TRY(launch_batch_finalize(g->t_plaq, n, g->d_plaq, s));
    HIP_TRY(g->ops->batch_polyakov_partial(g->t_poly, n, g->blocks_poly, s));
    HIP_TRY(launch_batch_finalize(g->t_poly, n, g->d_poly, s));
    HIP_TRY(launch_batch_record(g->t_plaq, n, g->d_plaq, g->d_poly, s));
    HIP_TRY(launch_batch_advance(g->t_update, n, s));
    return SUN_OK;
}

int capture_group(Group* g) {
    drop_graph(g);
    HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    const int rc = group_iteration(g);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != SUN_OK) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(SUN_ERR_HIP, "server: hipStreamEndCapture: %s", hipGetErrorString(e));
    g->graph = graph;
    HIP_TRY(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
    return SUN_OK;
}

// idle group: take in the admitted jobs, rebuild if needed, launch up to max_iters
// iterations (never more than a member still needs) and record the completion event
int launch_group(sun_server* S, Group* g, int max_iters) {
    if (!g->incoming.empty()) {
        for (Job* J : g->incoming) g->members.push_back(J);
        g->incoming.clear();
        g->dirty = true;
    }
    if (g->members.empty()) return SUN_OK;
    if (g->dirty) {
        SUN_TRY(rebuild_tables(g));
        SUN_TRY(capture_group(g));
        g->dirty = false;
    }
    int k = max_iters;
    for (const Job* J : g->members) k = std::min(k, J->d.n_iter - J->iters_done);
    // a member with no iteration left is retired at the completion that finished it and
    // resume refuses finished jobs, so k >= 1 here; were it not, launching would write
    // past that member's observable history
    if (k < 1) return SUN_OK;
    for (int i = 0; i < k; i++) HIP_TRY(hipGraphLaunch(g->exec, g->stream));
    HIP_TRY(hipEventRecord(g->done, g->stream));
    g->inflight = k;
    g->t_launch = now_since(S->t0);
    return SUN_OK;
}

Group* find_group(sun_server* S, const Job* J) {
    const int vk = group_vkey(J->lat);
    for (Group* g : S->groups)
        if (g->ops == J->lat->ops && g->n_or == J->d.n_or && g->reunit == J->d.reunit_every && g->vkey == vk)
            return g;
    return nullptr;
}

// admission, under S->mu
int admit(sun_server* S) {
    S->admit_needed = false;
    bool head_blocked = false;
    for (auto it = S->pending.begin(); it != S->pending.end();) {
        Job* J = *it;
        if ((int)S->running.size() >= S->max_jobs) break;
        const size_t off = S->arena.alloc(J->bytes);
        if (off == SIZE_MAX) {
            if (++J->passed_over > kMaxPassOver && it == S->pending.begin()) head_blocked = true;
            if (head_blocked) break;
            ++it;
            continue;
        }
        sun_lattice* L = nullptr;
        sun_lattice_desc d = {J->d.nc, {J->d.dims[0], J->d.dims[1], J->d.dims[2], J->d.dims[3]}, J->d.beta,
                              J->d.storage, J->d.precision};
        int rc = lattice_in_memory(S->ctx, &d, S->base + off, J->d.n_iter, &L);
        if (rc != SUN_OK) {
            S->arena.release(off);
            return rc;
        }
        J->off = off;
        J->lat = L;
        Group* g = find_group(S, J);
        if (!g) {
            rc = make_group(S, L, J->d.n_or, J->d.reunit_every, &g);
            if (rc != SUN_OK) {
                lattice_free(L);
                J->lat = nullptr;
                S->arena.release(off);
                return rc;
            }
            S->groups.push_back(g);
        }
        J->group = g;
        // queued behind whatever the group has in flight; touches only this job's memory
        hipError_t e;
        if (!J->resume_field.empty()) {
            e = hipMemcpyAsync(L->buf, J->resume_field.data(), J->resume_field.size(), hipMemcpyHostToDevice,
                               g->stream);
            if (e == hipSuccess && !J->resume_hist.empty())
                e = hipMemcpyAsync(L->hist, J->resume_hist.data(), sizeof(double2) * J->resume_hist.size(),
                                   hipMemcpyHostToDevice, g->stream);
        } else {
            e = J->d.hot_start ? L->ops->hot_start(L->f, L->g, J->d.seed, g->stream)
                               : L->ops->cold_start(L->f, 4 * L->g.V, g->stream);
        }
        if (e != hipSuccess) {
            hipGetLastError();
            return fail(SUN_ERR_HIP, "server: start of job %llu: %s", (unsigned long long)J->id,
                        hipGetErrorString(e));
        }
        g->incoming.push_back(J);
        J->state = SUN_JOB_RUNNING;
        J->t_start = now_since(S->t0);
        S->running.push_back(J);
        it = S->pending.erase(it);
    }
    return SUN_OK;
}

// Job file: "SUNJOB01", u32 version (1), u32 has_lattice, the job description field
// by field (i32 nc, i32 dims[4], f64 beta, i32 storage, i32 precision, u64 seed,
// i32 hot_start, i32 n_iter, i32 n_or, i32 reunit_every), i32 iters_done, then
// iters_done observable rows (4 f64: plaquette spatial/temporal, Polyakov re/im) and,
// with has_lattice, a native lattice checkpoint (seed, next sweep = iters_done).
constexpr char kJobMagic[8] = {'S', 'U', 'N', 'J', 'O', 'B', '0', '1'};

template <class T>
bool put(FILE* f, const T& v) {
    return fwrite(&v, sizeof(T), 1, f) == 1;
}
template <class T>
bool get(FILE* f, T& v) {
    return fread(&v, sizeof(T), 1, f) == 1;
}

// the lattice comes from the device (a job on it) or from the host bytes of a resumed
// job not yet admitted again
int write_job_file(const Job* J, const char* path, const std::vector<double2>& hist) {
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SUN_ERR_INVALID, "cannot open %s for writing", path);
    const sun_job_desc& d = J->d;
    const uint32_t version = 1, has_lattice = (J->lat || !J->resume_field.empty()) ? 1u : 0u;
    bool ok = fwrite(kJobMagic, 1, 8, f) == 8 && put(f, version) && put(f, has_lattice) && put(f, d.nc);
    for (int i = 0; i < 4; i++) ok = ok && put(f, d.dims[i]);
    ok = ok && put(f, d.beta) && put(f, d.storage) && put(f, d.precision) && put(f, d.seed) && put(f, d.hot_start) &&
         put(f, d.n_iter) && put(f, d.n_or) && put(f, d.reunit_every) && put(f, J->iters_done);
    ok = ok && fwrite(hist.data(), sizeof(double2), hist.size(), f) == hist.size();
    int rc = ok ? SUN_OK : fail(SUN_ERR_INVALID, "writing %s failed", path);
    if (rc == SUN_OK && J->lat) {
        rc = write_native(f, J->lat, d.seed, (uint32_t)J->iters_done, path);
    } else if (rc == SUN_OK && has_lattice) {
        if (!write_native_bytes(f, d, J->resume_field, d.seed, (uint32_t)J->iters_done))
            rc = fail(SUN_ERR_INVALID, "writing %s failed", path);
    }
    if (fclose(f) != 0 && rc == SUN_OK) rc = fail(SUN_ERR_INVALID, "writing %s failed", path);
    return rc;
}

void set_job_error(Job* J, int iteration, double2 plaq, double2 poly) {
    char buf[192];
    snprintf(buf, sizeof(buf), "iteration %d: plaquette (%g, %g), Polyakov loop (%g, %g) not finite or outside [-1, 1]",
             iteration, plaq.x, plaq.y, poly.x, poly.y);
    J->error = buf;
}

// under S->mu; the job's group is idle and the job is one of its members.  A job file
// that cannot be written fails the suspended job (its state and message say why)
// instead of stopping the scheduler.
int retire(sun_server* S, Job* J, bool done) {
    Group* g = J->group;
    bool failed = !J->error.empty();
    if (done && !failed) {
        J->hist.resize((size_t)2 * J->d.n_iter);
        HIP_TRY(hipMemcpy(J->hist.data(), J->lat->hist, sizeof(double2) * J->hist.size(), hipMemcpyDeviceToHost));
        for (int it = 0; it < J->d.n_iter && !failed; it++)
            if (!observables_in_range(J->hist[2 * it], J->hist[2 * it + 1])) {
                set_job_error(J, it, J->hist[2 * it], J->hist[2 * it + 1]);
                failed = true;
            }
    } else if (J->suspend && !failed) {
        std::vector<double2> rows((size_t)2 * J->iters_done);
        if (!rows.empty())
            HIP_TRY(hipMemcpy(rows.data(), J->lat->hist, sizeof(double2) * rows.size(), hipMemcpyDeviceToHost));
        if (write_job_file(J, J->suspend_path.c_str(), rows) != SUN_OK) {
            J->error = std::string("suspend: ") + sun_last_error();
            failed = true;
        }
    }
    g->members.erase(std::find(g->members.begin(), g->members.end(), J));
    g->dirty = true;
    lattice_free(J->lat);
    J->lat = nullptr;
    S->arena.release(J->off);
    J->off = SIZE_MAX;
    J->state = failed ? SUN_JOB_FAILED : done ? SUN_JOB_DONE : J->suspend ? SUN_JOB_SUSPENDED : SUN_JOB_CANCELLED;
    if (failed) J->hist.clear();
    J->t_end = now_since(S->t0);
    J->resume_field.clear();
    J->resume_field.shrink_to_fit();
    S->running.erase(std::find(S->running.begin(), S->running.end(), J));
    S->admit_needed = true;
    if (done && !failed) S->jobs_done++;
    return SUN_OK;
}

// a group's launched iterations have completed (its event has fired): the members'
// latest observables are checked, then finished, failed, cancelled and suspended
// members are retired
int complete_group(sun_server* S, Group* g) {
    const int k = g->inflight;
    g->inflight = 0;
    g->iter_ms = (now_since(S->t0) - g->t_launch) * 1e3 / k;
    const int n = (int)g->members.size();
    std::vector<double2> plaq(n), poly(n);
    HIP_TRY(hipMemcpyAsync(plaq.data(), g->d_plaq, sizeof(double2) * n, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(poly.data(), g->d_poly, sizeof(double2) * n, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    std::lock_guard<std::mutex> lk(S->mu);
    const std::vector<Job*> members = g->members;  // the order of the device tables
    for (int i = 0; i < n; i++) {
        Job* J = members[i];
        J->iters_done += k;
        S->link_updates += (uint64_t)k * (uint64_t)(1 + J->d.n_or) * 4ull * (uint64_t)J->lat->g.V;
        if (!observables_in_range(plaq[i], poly[i])) set_job_error(J, J->iters_done - 1, plaq[i], poly[i]);
        const bool done = J->iters_done >= J->d.n_iter;
        if (done || J->cancel || J->suspend || !J->error.empty()) SUN_TRY(retire(S, J, done));
    }
    S->steps++;
    S->cv.notify_all();
    return SUN_OK;
}

// cancelled jobs that have not joined their (idle) group's tables yet
int drop_cancelled_incoming(sun_server* S) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Group* g : S->groups) {
        if (g->inflight) continue;
        for (auto it = g->incoming.begin(); it != g->incoming.end();) {
            Job* J = *it;
            if (!J->cancel) {
                ++it;
                continue;
            }
            HIP_TRY(hipStreamSynchronize(g->stream));  // its start kernels
            it = g->incoming.erase(it);
            g->members.push_back(J);  // retire() takes it out of the members again
            SUN_TRY(retire(S, J, false));
        }
    }
    return SUN_OK;
}

// synchronous step: one iteration of every resident job
int step(sun_server* S) {
    const double t_begin = now_since(S->t0);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        SUN_TRY(admit(S));
    }
    SUN_TRY(drop_cancelled_incoming(S));
    std::vector<Group*> active;
    for (Group* g : S->groups) {
        SUN_TRY(launch_group(S, g, 1));
        if (g->inflight) active.push_back(g);
    }
    for (Group* g : active) HIP_TRY(hipEventSynchronize(g->done));
    for (Group* g : active) SUN_TRY(complete_group(S, g));
    std::lock_guard<std::mutex> lk(S->mu);
    S->busy_s += now_since(S->t0) - t_begin;
    return SUN_OK;
}

// the scheduling thread: groups progress independently
int worker_body(sun_server* S) {
    double busy_since = -1;
    for (;;) {
        bool stopping;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            bool any_inflight = false;
            for (Group* g : S->groups) any_inflight |= g->inflight > 0;
            if (!any_inflight) {
                if (busy_since >= 0) S->busy_s += now_since(S->t0) - busy_since;
                busy_since = -1;
                S->cv.wait(lk, [&] { return S->stop || S->admit_needed || !S->running.empty(); });
            }
            stopping = S->stop;
            if (!stopping && S->admit_needed) SUN_TRY(admit(S));
        }
        if (!stopping) SUN_TRY(drop_cancelled_incoming(S));
        bool progressed = false, any_inflight = false;
        for (Group* g : S->groups) {
            if (g->inflight) {
                const hipError_t q = hipEventQuery(g->done);
                if (q == hipErrorNotReady) {
                    any_inflight = true;
                    continue;
                }
                HIP_TRY(q);
                SUN_TRY(complete_group(S, g));
                progressed = true;
            }
            if (stopping) continue;
            const int iters = g->iter_ms > 0 ? (int)(kTargetMs / g->iter_ms) : 1;
            SUN_TRY(launch_group(S, g, std::min(std::max(iters, 1), kMaxItersPerLaunch)));
            if (g->inflight) {
                any_inflight = true;
                if (busy_since < 0) busy_since = now_since(S->t0);
            }
        }
        if (stopping && !any_inflight) {
            if (busy_since >= 0) {
                std::lock_guard<std::mutex> lk(S->mu);
                S->busy_s += now_since(S->t0) - busy_since;
            }
            return SUN_OK;
        }
        if (!progressed && any_inflight) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

void worker_loop(sun_server* S) {
    std::lock_guard<std::mutex> sl(S->step_mu);
    hipSetDevice(S->ctx->device);
    const int rc = worker_body(S);
    if (rc != SUN_OK) {
        // drain what is in flight so no kernel outlives the failure
        for (Group* g : S->groups)
            if (g->inflight) hipEventSynchronize(g->done);
        std::lock_guard<std::mutex> lk(S->mu);
        S->worker_rc = rc;
        S->stop = true;
        S->cv.notify_all();
    }
}

Job* find_job(sun_server* S, uint64_t id) {
    auto it = S->jobs.find(id);
    return it == S->jobs.end() ? nullptr : it->second;
}

}  // namespace

extern "C" int sun_server_create(sun_ctx* ctx, size_t budget, int max_jobs, sun_server** out) {
    if (!ctx || !out || budget < 1024 * 1024 || max_jobs < 1 || max_jobs > 1024)
        return fail(SUN_ERR_INVALID, "sun_server_create: bad argument (budget >= 1 MiB, 1 <= max_jobs <= 1024)");
    *out = nullptr;
    sun_server* S = new (std::nothrow) sun_server();
    if (!S) return fail(SUN_ERR_NOMEM, "host allocation");
    S->ctx = ctx;
    S->budget = budget;
    S->max_jobs = max_jobs;
    HIP_TRY_CLEANUP(hipSetDevice(ctx->device), delete S);
    const hipError_t e = hipMalloc(&S->base, budget);
    if (e != hipSuccess) {
        hipGetLastError();
        delete S;
        return fail(SUN_ERR_NOMEM, "sun_server_create: reserving %zu bytes: %s", budget, hipGetErrorString(e));
    }
    S->arena.reset(budget);
    S->t0 = std::chrono::steady_clock::now();
    *out = S;
    return SUN_OK;
}

extern "C" int sun_server_stop(sun_server* S) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_stop: NULL");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    std::lock_guard<std::mutex> lk(S->mu);
    S->worker_on = false;
    S->stop = false;
    const int rc = S->worker_rc;
    S->worker_rc = SUN_OK;
    return rc == SUN_OK ? SUN_OK : fail(rc, "the scheduling thread had stopped on an error");
}

extern "C" int sun_server_destroy(sun_server* S) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_destroy: NULL");
    sun_server_stop(S);
    hipSetDevice(S->ctx->device);
    for (Group* g : S->groups) {
        hipStreamSynchronize(g->stream);
        free_group(g);
    }
    for (auto& kv : S->jobs) {
        if (kv.second->lat) lattice_free(kv.second->lat);
        delete kv.second;
    }
    hipFree(S->base);
    delete S;
    return SUN_OK;
}

// checks of a job description (submitted, or read from a job file: with_state when
// the file carries the job's lattice) and the server memory it needs; host only
int check_job_desc(const sun_job_desc* jd, bool with_state, size_t* bytes) {
    sun_lattice_desc d = {jd->nc, {jd->dims[0], jd->dims[1], jd->dims[2], jd->dims[3]}, jd->beta, jd->storage,
                          jd->precision};
    SUN_TRY(check_desc(&d));
    if (jd->n_iter < 1 || jd->n_or < 0) return fail(SUN_ERR_INVALID, "job: n_iter >= 1 and n_or >= 0 required");
    if (jd->reunit_every != 0 && jd->reunit_every != 1)
        return fail(SUN_ERR_UNSUPPORTED, "job: reunit_every must be 0 or 1 in the server");
    if (jd->storage == SUN_STORE_12 && jd->reunit_every == 0)
        return fail(SUN_ERR_UNSUPPORTED,
                    "job: 12-real links need reunit_every = 1 (row 2 is rebuilt from rows 0 and 1, whose drift "
                    "from SU(3) grows about 4x per iteration without reunitarisation)");
    if (!with_state && !jd->hot_start && jd->storage == SUN_STORE_8)
        return fail(SUN_ERR_UNSUPPORTED, "job: 8-real storage cannot hold a cold start (gaugearray.cu:293-297)");
    *bytes = Arena::round_up(lattice_footprint(&d, jd->n_iter));
    return SUN_OK;
}

int check_job(const sun_server* S, const sun_job_desc* jd, bool with_state, size_t* bytes) {
    SUN_TRY(check_job_desc(jd, with_state, bytes));
    if (*bytes > S->arena.capacity())
        return fail(SUN_ERR_BUDGET, "job needs %zu bytes, the server budget is %zu", *bytes, S->arena.capacity());
    return SUN_OK;
}

extern "C" int sun_job_bytes(const sun_job_desc* jd, size_t* bytes) {
    if (!jd || !bytes) return fail(SUN_ERR_INVALID, "sun_job_bytes: NULL argument");
    return check_job_desc(jd, false, bytes);
}

extern "C" int sun_server_submit(sun_server* S, const sun_job_desc* jd, uint64_t* id) {
    if (!S || !jd || !id) return fail(SUN_ERR_INVALID, "sun_server_submit: NULL argument");
    size_t bytes = 0;
    SUN_TRY(check_job(S, jd, false, &bytes));
    Job* J = new (std::nothrow) Job();
    if (!J) return fail(SUN_ERR_NOMEM, "host allocation");
    J->d = *jd;
    J->bytes = bytes;
    std::lock_guard<std::mutex> lk(S->mu);
    J->id = S->next_id++;
    J->t_submit = now_since(S->t0);
    S->jobs[J->id] = J;
    S->pending.push_back(J);
    S->admit_needed = true;
    *id = J->id;
    S->cv.notify_all();
    return SUN_OK;
}

extern "C" int sun_server_cancel(sun_server* S, uint64_t id) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_cancel: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    if (finished(J)) return SUN_OK;
    J->cancel = true;
    if (J->state == SUN_JOB_PENDING) {
        S->pending.erase(std::find(S->pending.begin(), S->pending.end(), J));
        J->state = SUN_JOB_CANCELLED;
        J->t_end = now_since(S->t0);
        S->admit_needed = true;
    }
    S->cv.notify_all();
    return SUN_OK;
}

extern "C" int sun_server_step(sun_server* S) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_step: NULL");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->worker_on) return fail(SUN_ERR_STATE, "sun_server_step: the scheduling thread is running");
    }
    std::lock_guard<std::mutex> sl(S->step_mu);
    HIP_TRY(hipSetDevice(S->ctx->device));
    return step(S);
}

extern "C" int sun_server_start(sun_server* S) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_start: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->worker_on) return SUN_OK;
    S->stop = false;
    S->worker_rc = SUN_OK;
    S->worker_on = true;
    S->admit_needed = true;
    S->worker = std::thread(worker_loop, S);
    return SUN_OK;
}

extern "C" int sun_server_wait(sun_server* S, uint64_t id, double timeout_s) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_wait: NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    auto ended = [&] { return finished(J) || S->worker_rc != SUN_OK; };
    if (!S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ended))
        return fail(SUN_ERR_STATE, "job %llu not finished after %.1f s", (unsigned long long)id, timeout_s);
    if (!finished(J))
        return fail(SUN_ERR_STATE, "the scheduling thread stopped on an error before job %llu finished",
                    (unsigned long long)id);
    return SUN_OK;
}

extern "C" int sun_server_status(sun_server* S, uint64_t id, sun_job_status* st) {
    if (!S || !st) return fail(SUN_ERR_INVALID, "sun_server_status: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    const Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    memset(st, 0, sizeof(*st));
    st->id = J->id;
    st->state = J->state;
    st->iters_done = J->iters_done;
    st->t_submit = J->t_submit;
    st->t_start = J->t_start;
    st->t_end = J->t_end;
    if (J->state == SUN_JOB_DONE && !J->hist.empty()) {
        const size_t last = J->hist.size() - 2;
        st->plaq[0] = J->hist[last].x;
        st->plaq[1] = J->hist[last].y;
        st->poly[0] = J->hist[last + 1].x;
        st->poly[1] = J->hist[last + 1].y;
    }
    return SUN_OK;
}

extern "C" int sun_server_observables(sun_server* S, uint64_t id, double* out, int max_rows) {
    if (!S || !out) return fail(SUN_ERR_INVALID, "sun_server_observables: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    const Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    if (J->state != SUN_JOB_DONE) return fail(SUN_ERR_STATE, "job %llu is not done", (unsigned long long)id);
    const int rows = std::min(max_rows, J->d.n_iter);
    for (int r = 0; r < rows; r++) {
        out[4 * r + 0] = J->hist[2 * r].x;
        out[4 * r + 1] = J->hist[2 * r].y;
        out[4 * r + 2] = J->hist[2 * r + 1].x;
        out[4 * r + 3] = J->hist[2 * r + 1].y;
    }
    return SUN_OK;
}

extern "C" int sun_server_release(sun_server* S, uint64_t id) {
    if (!S) return fail(SUN_ERR_INVALID, "sun_server_release: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    if (!finished(J)) return fail(SUN_ERR_STATE, "job %llu is still pending or running", (unsigned long long)id);
    S->jobs.erase(id);
    delete J;
    return SUN_OK;
}

extern "C" int sun_server_stats(sun_server* S, sun_server_counters* st) {
    if (!S || !st) return fail(SUN_ERR_INVALID, "sun_server_stats: NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    st->jobs_done = S->jobs_done;
    st->link_updates = S->link_updates;
    st->steps = S->steps;
    st->busy_s = S->busy_s;
    st->budget = S->arena.capacity();
    st->bytes_in_use = S->arena.in_use();
    st->running = (int)S->running.size();
    st->pending = (int)S->pending.size();
    return SUN_OK;
}

extern "C" double sun_server_clock(sun_server* S) { return S ? now_since(S->t0) : 0.0; }

extern "C" int sun_server_suspend(sun_server* S, uint64_t id, const char* path) {
    if (!S || !path) return fail(SUN_ERR_INVALID, "sun_server_suspend: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    if (J->state != SUN_JOB_PENDING && J->state != SUN_JOB_RUNNING)
        return fail(SUN_ERR_STATE, "job %llu is not pending or running", (unsigned long long)id);
    if (J->suspend) return fail(SUN_ERR_STATE, "job %llu is already being suspended", (unsigned long long)id);
    J->suspend_path = path;
    if (J->state == SUN_JOB_PENDING) {
        // not on the device: the file holds the description and any resumed state
        SUN_TRY(write_job_file(J, path, J->resume_hist));
        S->pending.erase(std::find(S->pending.begin(), S->pending.end(), J));
        J->suspend = true;
        J->state = SUN_JOB_SUSPENDED;
        J->t_end = now_since(S->t0);
        J->resume_field.clear();
        S->admit_needed = true;
        S->cv.notify_all();
        return SUN_OK;
    }
    // the file is written at the group's next completion; a path that cannot be
    // created is refused now, while the job can still run on
    FILE* probe = fopen(path, "wb");
    if (!probe) return fail(SUN_ERR_INVALID, "cannot open %s for writing", path);
    fclose(probe);
    J->suspend = true;  // retired, with its file written, at the group's next completion
    S->cv.notify_all();
    return SUN_OK;
}

extern "C" int sun_server_job_error(sun_server* S, uint64_t id, char* buf, int len) {
    if (!S || !buf || len < 1) return fail(SUN_ERR_INVALID, "sun_server_job_error: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    const Job* J = find_job(S, id);
    if (!J) return fail(SUN_ERR_STATE, "unknown job %llu", (unsigned long long)id);
    snprintf(buf, (size_t)len, "%s", J->error.c_str());
    return SUN_OK;
}

extern "C" int sun_server_resume(sun_server* S, const char* path, uint64_t* id) {
    if (!S || !path || !id) return fail(SUN_ERR_INVALID, "sun_server_resume: NULL argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(SUN_ERR_INVALID, "cannot open %s", path);
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{f};
    char magic[8];
    uint32_t version = 0, has_lattice = 0;
    sun_job_desc d;
    memset(&d, 0, sizeof(d));
    int iters_done = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kJobMagic, 8) == 0 && get(f, version) && version == 1 &&
              get(f, has_lattice) && get(f, d.nc);
    for (int i = 0; i < 4; i++) ok = ok && get(f, d.dims[i]);
    ok = ok && get(f, d.beta) && get(f, d.storage) && get(f, d.precision) && get(f, d.seed) && get(f, d.hot_start) &&
         get(f, d.n_iter) && get(f, d.n_or) && get(f, d.reunit_every) && get(f, iters_done);
    if (!ok) return fail(SUN_ERR_INVALID, "%s: not a job file", path);
    sun_lattice_desc ld = {d.nc, {d.dims[0], d.dims[1], d.dims[2], d.dims[3]}, d.beta, d.storage, d.precision};
    // a job file never holds a finished job (a job that completes is retired as done,
    // not suspended), so iters_done < n_iter
    if (iters_done < 0 || iters_done >= d.n_iter || (iters_done > 0 && !has_lattice) || has_lattice > 1)
        return fail(SUN_ERR_INVALID, "%s: inconsistent job state", path);
    // every size below follows from a description that fits the budget
    size_t bytes = 0;
    SUN_TRY(check_job(S, &d, has_lattice != 0, &bytes));
    Job* J = new (std::nothrow) Job();
    if (!J) return fail(SUN_ERR_NOMEM, "host allocation");
    J->d = d;
    J->iters_done = iters_done;
    J->bytes = bytes;
    try {
        J->resume_hist.resize((size_t)2 * iters_done);
    } catch (const std::bad_alloc&) {
        delete J;
        return fail(SUN_ERR_NOMEM, "%s: host allocation for %d observable rows", path, iters_done);
    }
    if (fread(J->resume_hist.data(), sizeof(double2), J->resume_hist.size(), f) != J->resume_hist.size()) {
        delete J;
        return fail(SUN_ERR_INVALID, "%s: truncated observables", path);
    }
    if (has_lattice) {
        uint64_t seed = 0;
        uint32_t next = 0;
        const int rc = read_native(f, ld, 4 * (size_t)make_geom(ld.dims).V * link_bytes(&ld), J->resume_field, &seed,
                                   &next, path);
        if (rc != SUN_OK || seed != d.seed || (int)next != iters_done) {
            delete J;
            return rc != SUN_OK ? rc : fail(SUN_ERR_INVALID, "%s: lattice checkpoint does not match the job", path);
        }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    J->id = S->next_id++;
    J->t_submit = now_since(S->t0);
    S->jobs[J->id] = J;
    S->pending.push_back(J);
    S->admit_needed = true;
    *id = J->id;
    S->cv.notify_all();
    return SUN_OK;
}
