// sun_generate.hip -- the ensemble-generation driver for one lattice and for a
// batch of lattices.  One iteration = heatbath sweep, n_or overrelaxation sweeps,
// optional reunitarisation, plaquette (+ Polyakov loop for a single lattice).  With
// use_graph the iteration is captured once into a hipGraph and replayed; the
// heatbath sweep number and the observable slot live in device counters that the
// graph itself advances, so every replay continues the Markov chain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include "sun_internal.hpp"

using namespace sun;

namespace {

struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphGuard() {
        if (x) hipGraphExecDestroy(x);
        if (g) hipGraphDestroy(g);
    }
};

int ensure_obs(double2** buf, int* cap, int n) {
    if (*cap >= n) return SUN_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc(buf, sizeof(double2) * (size_t)n));
    *cap = n;
    return SUN_OK;
}

// one iteration of a single lattice; sweep taken from lat->sweep_dev
int lattice_iteration(sun_lattice* lat, uint64_t seed, int n_or, bool reunit, bool measure, uint32_t first,
                      hipStream_t s) {
    SUN_TRY(heatbath_sweep(lat, seed, 0, lat->sweep_dev, s));
    for (int k = 0; k < n_or; k++) SUN_TRY(overrelax_sweep(lat, s));
    if (reunit) HIP_TRY(lat->ops->reunitarize(lat->f, 4 * lat->g.V, s));
    if (measure) {
        // plaquette and Polyakov loop in one grid, then one workgroup that finalizes
        // both (the sums of sun_plaquette / sun_polyakov, bit for bit), records them
        // and advances the sweep counter
        const double ts = (double)lat->g.X[0] * lat->g.X[1] * lat->g.X[2];
        HIP_TRY(lat->ops->measure_partial(lat->f, lat->g, lat->partials, s));
        HIP_TRY(launch_finalize_iteration(lat->partials, plaquette_blocks(lat->g), polyakov_blocks(lat->g),
                                          1.0 / ((double)(3 * lat->d.nc) * (double)lat->g.V),
                                          1.0 / ((double)lat->d.nc * ts), lat->dres, lat->obs_dev, lat->sweep_dev,
                                          first, s));
    } else {
        HIP_TRY(launch_advance_counter(lat->sweep_dev, s));
    }
    return SUN_OK;
}

int capture(hipStream_t s, GraphGuard& gg, const std::function<int()>& body) {
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = body();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != SUN_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(SUN_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    gg.g = g;
    HIP_TRY(hipGraphInstantiate(&gg.x, g, nullptr, nullptr, 0));
    return SUN_OK;
}

}  // namespace

static int sun_generate_body(sun_lattice* lat, const sun_generate_desc* d, void* stream, double* obs) {
    if (!lat || !d) return fail(SUN_ERR_INVALID, "sun_generate: NULL argument");
    if (d->n_iter < 0 || d->n_or < 0 || d->reunit_every < 0)
        return fail(SUN_ERR_INVALID, "sun_generate: negative count");
    if (lat->d.storage == SUN_STORE_12 && d->reunit_every == 0)
        return fail(SUN_ERR_UNSUPPORTED,
                    "sun_generate: 12-real links need reunit_every >= 1 (row 2 is rebuilt from rows 0 and 1, whose "
                    "drift from SU(3) grows about 4x per iteration without reunitarisation)");
    if (d->n_iter == 0) return SUN_OK;
    hipStream_t s = pick_stream(lat, stream);
    const bool measure = d->measure != 0;
    if (measure) SUN_TRY(ensure_obs(&lat->obs_dev, &lat->obs_cap, 2 * d->n_iter));
    const uint32_t first = d->first_sweep;
    SUN_TRY(tune_lattice(lat, s));
    HIP_TRY(hipMemcpyAsync(lat->sweep_dev, &first, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));

    if (d->use_graph) {
        GraphGuard plain, with_reunit;
        const bool need_plain = d->reunit_every != 1;
        const bool need_reunit = d->reunit_every >= 1;
        if (need_plain)
            SUN_TRY(capture(s, plain, [&] { return lattice_iteration(lat, d->seed, d->n_or, false, measure, first, s); }));
        if (need_reunit)
            SUN_TRY(capture(s, with_reunit,
                            [&] { return lattice_iteration(lat, d->seed, d->n_or, true, measure, first, s); }));
        for (int it = 0; it < d->n_iter; it++) {
            const bool ru = d->reunit_every > 0 && ((it + 1) % d->reunit_every) == 0;
            HIP_TRY(hipGraphLaunch(ru ? with_reunit.x : plain.x, s));
        }
    } else {
        for (int it = 0; it < d->n_iter; it++) {
            const bool ru = d->reunit_every > 0 && ((it + 1) % d->reunit_every) == 0;
            SUN_TRY(lattice_iteration(lat, d->seed, d->n_or, ru, measure, first, s));
        }
    }
    if (measure && obs) {
        std::vector<double2> h((size_t)2 * d->n_iter);
        HIP_TRY(hipMemcpyAsync(h.data(), lat->obs_dev, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int it = 0; it < d->n_iter; it++) {
            obs[4 * it + 0] = h[2 * it].x;
            obs[4 * it + 1] = h[2 * it].y;
            obs[4 * it + 2] = h[2 * it + 1].x;
            obs[4 * it + 3] = h[2 * it + 1].y;
        }
        // a field that has overflowed or left SU(N) (the reference stops the process,
        // errorCULQCD, comm_mpi.h:116-120): the rows are returned and the call fails
        for (int it = 0; it < d->n_iter; it++)
            if (!observables_in_range(h[2 * it], h[2 * it + 1]))
                return fail(SUN_ERR_STATE,
                            "sun_generate: iteration %d: plaquette (%g, %g), Polyakov loop (%g, %g) not finite or "
                            "outside [-1, 1]",
                            it, h[2 * it].x, h[2 * it].y, h[2 * it + 1].x, h[2 * it + 1].y);
    } else {
        HIP_TRY(hipStreamSynchronize(s));
    }
    return SUN_OK;
}

extern "C" int sun_generate(sun_lattice* lat, const sun_generate_desc* d, void* stream, double* obs) {
    return guard_alloc("sun_generate", [&] { return sun_generate_body(lat, d, stream, obs); });
}

// ---------------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------------
static void batch_free(sun_batch* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    for (sun_batch::Schedule* sc : {&b->gen, &b->step}) {
        if (sc->t_update) hipFree(sc->t_update);
        if (sc->t_reunit) hipFree(sc->t_reunit);
        if (sc->t_plaq) hipFree(sc->t_plaq);
    }
    if (b->partials) hipFree(b->partials);
    if (b->d_out) hipFree(b->d_out);
    if (b->h_out) hipHostFree(b->h_out);
    if (b->obs_dev) hipFree(b->obs_dev);
    if (b->iter_dev) hipFree(b->iter_dev);
    for (int k = 1; k < kBatchBranches; k++) {
        if (b->branch[k]) hipStreamDestroy(b->branch[k]);
        if (b->join[k]) hipEventDestroy(b->join[k]);
    }
    if (b->fork) hipEventDestroy(b->fork);
    delete b;
}

// the chunks and tables of one schedule: chunks of consecutive lattices within budget
// link bytes; entries with first_block counted from their chunk's start
static void build_schedule(sun_batch::Schedule& sc, sun_lattice* const* lats, int n, const uint64_t* seeds,
                           size_t budget, int branches, std::vector<BatchEntry>& tr, std::vector<BatchEntry>& tp) {
    std::vector<BatchEntry> tu(n);
    tr.assign(n, BatchEntry{});
    tp.assign(n, BatchEntry{});
    sc.branches = branches;
    int cu = 0, cr = 0, cp = 0;
    size_t chunk_bytes = 0;
    for (int i = 0; i < n; i++) {
        sun_lattice* L = lats[i];
        if (sc.chunks.empty() || (sc.chunks.back().n > 0 && chunk_bytes + L->bytes > budget)) {
            sc.chunks.push_back(sun_batch::Chunk{i, 0, 0, 0, 0});
            chunk_bytes = 0;
            cu = cr = cp = 0;
        }
        chunk_bytes += L->bytes;
        BatchEntry e;
        memset(&e, 0, sizeof(e));
        e.f = L->f;
        e.g = L->g;
        e.seed = seeds ? seeds[i] : 0;
        e.bnc = L->d.beta / L->d.nc;
        e.sweep = 0;
        batch_brick(L->g, e.ub1, e.ub2, e.ub3, e.spt);
        BatchEntry a = e, r = e, p = e;
        sun_batch::Chunk& c = sc.chunks.back();
        c.n++;
        a.first_block = cu;
        a.nblocks = update_blocks(L->g, 256);
        cu += a.nblocks;
        c.blocks = cu;
        r.first_block = cr;
        r.nblocks = link_blocks(L->g);
        cr += r.nblocks;
        c.blocks_reunit = cr;
        p.first_block = cp;
        p.nblocks = plaquette_blocks(L->g);
        p.scale = 1.0 / ((double)(3 * L->d.nc) * (double)L->g.V);  // plaquette_cub.cu:234
        cp += p.nblocks;
        c.blocks_plaq = cp;
        tu[i] = a;
        tr[i] = r;
        tp[i] = p;
    }
    sc.h_update = tu;
}

static hipError_t upload_schedule(sun_batch::Schedule& sc, const std::vector<BatchEntry>& tr,
                                  const std::vector<BatchEntry>& tp) {
    const size_t bytes = sizeof(BatchEntry) * sc.h_update.size();
    hipError_t e = hipMalloc(&sc.t_update, bytes);
    if (e == hipSuccess) e = hipMalloc(&sc.t_reunit, bytes);
    if (e == hipSuccess) e = hipMalloc(&sc.t_plaq, bytes);
    if (e == hipSuccess) e = hipMemcpy(sc.t_update, sc.h_update.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sc.t_reunit, tr.data(), bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sc.t_plaq, tp.data(), bytes, hipMemcpyHostToDevice);
    return e;
}

static int sun_batch_create_body(sun_ctx* ctx, int n, sun_lattice* const* lats, const uint64_t* seeds,
                                 sun_batch** out, sun_batch*& b);

extern "C" int sun_batch_create(sun_ctx* ctx, int n, sun_lattice* const* lats, const uint64_t* seeds,
                                sun_batch** out) {
    sun_batch* b = nullptr;  // freed here when a host allocation fails part way
    const int rc = guard_alloc("sun_batch_create", [&] { return sun_batch_create_body(ctx, n, lats, seeds, out, b); });
    if (rc == SUN_ERR_NOMEM && b && out && *out != b) batch_free(b);
    return rc;
}

static int sun_batch_create_body(sun_ctx* ctx, int n, sun_lattice* const* lats, const uint64_t* seeds,
                                 sun_batch** out, sun_batch*& b) {
    if (!ctx || !lats || !out || n <= 0 || n > 1024)
        return fail(SUN_ERR_INVALID, "sun_batch_create: bad argument (n=%d, 1..1024)", n);
    *out = nullptr;
    const OpsTable* ops = lats[0] ? lats[0]->ops : nullptr;
    for (int i = 0; i < n; i++) {
        if (!lats[i]) return fail(SUN_ERR_INVALID, "sun_batch_create: lattice %d is NULL", i);
        if (lats[i]->ops != ops)
            return fail(SUN_ERR_INVALID, "sun_batch_create: lattice %d differs in N/storage/precision", i);
        if (lats[i]->ctx != ctx) return fail(SUN_ERR_INVALID, "sun_batch_create: lattice %d has another context", i);
    }
    b = new (std::nothrow) sun_batch();
    if (!b) return fail(SUN_ERR_NOMEM, "host allocation");
    b->ctx = ctx;
    b->ops = ops;
    for (int k = 0; k < 2; k++) {
        b->piped[k] = default_batch_piped(ops, k);
        b->brick[k] = default_batch_brick(ops, k);
        b->lds[k] = default_batch_lds(ops, k);
    }
    b->lats.assign(lats, lats + n);
    // each lattice's plaquette partials in b->partials (the same in both schedules)
    std::vector<int> poff(n);
    int cp_all = 0;
    for (int i = 0; i < n; i++) {
        poff[i] = cp_all;
        cp_all += plaquette_blocks(lats[i]->g);
    }
    std::vector<BatchEntry> tr_gen, tp_gen, tr_step, tp_step;
    build_schedule(b->gen, lats, n, seeds, kBatchChunkBytes, kBatchBranches, tr_gen, tp_gen);
    build_schedule(b->step, lats, n, seeds, kCacheChunkBytes, 1, tr_step, tp_step);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(&b->partials, sizeof(double2) * (size_t)cp_all);
    if (e == hipSuccess) e = hipMalloc(&b->d_out, sizeof(double2) * n);
    if (e == hipSuccess)
        for (int i = 0; i < n; i++) tp_gen[i].partials = tp_step[i].partials = b->partials + poff[i];
    if (e == hipSuccess) e = hipMalloc(&b->iter_dev, sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(&b->h_out, sizeof(double2) * n, hipHostMallocDefault);
    if (e == hipSuccess) e = upload_schedule(b->gen, tr_gen, tp_gen);
    if (e == hipSuccess) e = upload_schedule(b->step, tr_step, tp_step);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (int k = 1; k < kBatchBranches && e == hipSuccess; k++) {
        e = hipStreamCreateWithFlags(&b->branch[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        hipGetLastError();
        batch_free(b);
        b = nullptr;
        return fail(SUN_ERR_NOMEM, "sun_batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return SUN_OK;
}

extern "C" int sun_batch_destroy(sun_batch* b) {
    if (!b) return fail(SUN_ERR_INVALID, "sun_batch_destroy: NULL");
    hipSetDevice(b->ctx->device);
    hipDeviceSynchronize();
    batch_free(b);
    return SUN_OK;
}

extern "C" int sun_batch_get_launch(const sun_batch* b, int kernel, int* piped, int* brick, int* lds) {
    if (!b || !piped || !brick || !lds || (kernel != 0 && kernel != 1))
        return fail(SUN_ERR_INVALID, "sun_batch_get_launch: bad argument");
    *piped = b->piped[kernel];
    *brick = b->brick[kernel];
    *lds = b->lds[kernel];
    return SUN_OK;
}

extern "C" int sun_batch_set_launch(sun_batch* b, int kernel, int piped, int brick, int lds) {
    if (!b || (kernel != 0 && kernel != 1)) return fail(SUN_ERR_INVALID, "sun_batch_set_launch: bad argument");
    if (piped < 0 || piped > 2) return fail(SUN_ERR_INVALID, "sun_batch_set_launch: piped %d (0, 1 or 2)", piped);
    if (brick != 0 && brick != 1) return fail(SUN_ERR_INVALID, "sun_batch_set_launch: brick %d (0 or 1)", brick);
    if (lds < 0 || lds > 65536) return fail(SUN_ERR_INVALID, "sun_batch_set_launch: lds %d (0..65536)", lds);
    if (piped && !b->ops->has_piped)
        return fail(SUN_ERR_INVALID, "sun_batch_set_launch: no pipelined kernel for this N/storage/precision");
    b->piped[kernel] = piped;
    b->brick[kernel] = brick;
    b->lds[kernel] = lds;
    return SUN_OK;
}

static hipStream_t batch_stream(sun_batch* b, void* stream) {
    return stream ? (hipStream_t)stream : b->ctx->stream;
}

static int batch_set_sweeps(sun_batch* b, sun_batch::Schedule& sc, const uint32_t* sweeps, uint32_t uniform,
                            hipStream_t s) {
    const int n = (int)b->lats.size();
    for (int i = 0; i < n; i++) sc.h_update[i].sweep = sweeps ? sweeps[i] : uniform;
    HIP_TRY(hipMemcpyAsync(sc.t_update, sc.h_update.data(), sizeof(BatchEntry) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

using Schedule = sun_batch::Schedule;
using Chunk = sun_batch::Chunk;

// one heatbath (hb) or overrelaxation sweep of one update chunk
static int chunk_sweep(sun_batch* b, const Schedule& sc, const Chunk& c, bool hb, hipStream_t s) {
    const int k = hb ? 0 : 1;
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++)
            HIP_TRY(b->ops->batch_update(sc.t_update + c.first, c.n, c.blocks, parity, mu, hb, b->piped[k],
                                         b->brick[k], b->lds[k], s));
    return SUN_OK;
}

static int chunk_reunit(sun_batch* b, const Schedule& sc, const Chunk& c, hipStream_t s) {
    HIP_TRY(b->ops->batch_reunitarize(sc.t_reunit + c.first, c.n, c.blocks_reunit, s));
    return SUN_OK;
}

static int chunk_plaq(sun_batch* b, const Schedule& sc, const Chunk& c, hipStream_t s) {
    HIP_TRY(b->ops->batch_plaquette_partial(sc.t_plaq + c.first, c.n, c.blocks_plaq, s));
    HIP_TRY(launch_batch_finalize(sc.t_plaq + c.first, c.n, b->d_out + c.first, s));
    return SUN_OK;
}

// body(chunk, stream) for every chunk of the schedule, chunk i on branch
// i % sc.branches: branch 0 is s itself; the other branches first wait for the work
// already on s, and s waits for them at the end, so to the caller everything ran on s.
// A chunk's lattices are in no other chunk, so each lattice's launches keep their order.
template <typename Body>
static int for_chunks(sun_batch* b, const Schedule& sc, hipStream_t s, Body&& body) {
    const int nb = std::min<int>(sc.branches, (int)sc.chunks.size());
    if (nb > 1) {
        HIP_TRY(hipEventRecord(b->fork, s));
        for (int k = 1; k < nb; k++) HIP_TRY(hipStreamWaitEvent(b->branch[k], b->fork, 0));
    }
    int rc = SUN_OK;
    for (size_t i = 0; i < sc.chunks.size() && rc == SUN_OK; i++) {
        const int k = (int)(i % (size_t)nb);
        rc = body(sc.chunks[i], k == 0 ? s : b->branch[k]);
    }
    // joined also after a failed launch, so no branch is left outside s (or outside a
    // capture begun on s)
    for (int k = 1; k < nb; k++) {
        const hipError_t e1 = hipEventRecord(b->join[k], b->branch[k]);
        const hipError_t e2 = hipStreamWaitEvent(s, b->join[k], 0);
        if (rc == SUN_OK && (e1 != hipSuccess || e2 != hipSuccess))
            rc = fail(SUN_ERR_HIP, "batch branch join: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    return rc;
}

extern "C" int sun_batch_heatbath(sun_batch* b, const uint32_t* sweeps, void* stream) {
    if (!b || !sweeps) return fail(SUN_ERR_INVALID, "sun_batch_heatbath: NULL argument");
    hipStream_t s = batch_stream(b, stream);
    SUN_TRY(batch_set_sweeps(b, b->step, sweeps, 0, s));
    return for_chunks(b, b->step, s, [&](const Chunk& c, hipStream_t cs) { return chunk_sweep(b, b->step, c, true, cs); });
}

extern "C" int sun_batch_overrelax(sun_batch* b, void* stream) {
    if (!b) return fail(SUN_ERR_INVALID, "sun_batch_overrelax: NULL");
    return for_chunks(b, b->step, batch_stream(b, stream),
                      [&](const Chunk& c, hipStream_t cs) { return chunk_sweep(b, b->step, c, false, cs); });
}

extern "C" int sun_batch_reunitarize(sun_batch* b, void* stream) {
    if (!b) return fail(SUN_ERR_INVALID, "sun_batch_reunitarize: NULL");
    return for_chunks(b, b->step, batch_stream(b, stream),
                      [&](const Chunk& c, hipStream_t cs) { return chunk_reunit(b, b->step, c, cs); });
}

extern "C" int sun_batch_plaquette(sun_batch* b, void* stream, double* out) {
    if (!b || !out) return fail(SUN_ERR_INVALID, "sun_batch_plaquette: NULL");
    hipStream_t s = batch_stream(b, stream);
    const int n = (int)b->lats.size();
    SUN_TRY(for_chunks(b, b->step, s, [&](const Chunk& c, hipStream_t cs) { return chunk_plaq(b, b->step, c, cs); }));
    HIP_TRY(hipMemcpyAsync(b->h_out, b->d_out, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) {
        out[2 * i] = b->h_out[i].x;
        out[2 * i + 1] = b->h_out[i].y;
    }
    return SUN_OK;
}

// one iteration: chunk by chunk (over the branches of for_chunks), the heatbath sweep,
// n_or overrelaxation sweeps, reunitarisation and plaquette back to back (lattices are
// independent, so every lattice sees the same launches in the same order as without
// chunks); then the history row of the whole batch
static int batch_iteration(sun_batch* b, int n_or, bool reunit, bool measure, hipStream_t s) {
    const int n = (int)b->lats.size();
    const Schedule& sc = b->gen;
    SUN_TRY(for_chunks(b, sc, s, [&](const Chunk& c, hipStream_t cs) {
        SUN_TRY(chunk_sweep(b, sc, c, true, cs));
        for (int k = 0; k < n_or; k++) SUN_TRY(chunk_sweep(b, sc, c, false, cs));
        if (reunit) SUN_TRY(chunk_reunit(b, sc, c, cs));
        if (measure) SUN_TRY(chunk_plaq(b, sc, c, cs));
        return SUN_OK;
    }));
    if (measure) HIP_TRY(launch_record_obs(b->d_out, n, b->obs_dev, b->iter_dev, 0, s));
    HIP_TRY(launch_batch_advance(sc.t_update, n, s));
    HIP_TRY(launch_advance_counter(b->iter_dev, s));
    return SUN_OK;
}

static int sun_batch_generate_body(sun_batch* b, uint32_t first_sweep, int n_iter, int n_or, int reunit_every,
                                   int use_graph, void* stream, double* obs) {
    if (!b || n_iter < 0 || n_or < 0 || reunit_every < 0) return fail(SUN_ERR_INVALID, "sun_batch_generate: bad argument");
    if (n_iter == 0) return SUN_OK;
    hipStream_t s = batch_stream(b, stream);
    const int n = (int)b->lats.size();
    const bool measure = obs != nullptr;
    if (measure) SUN_TRY(ensure_obs(&b->obs_dev, &b->obs_cap, n * n_iter));
    SUN_TRY(batch_set_sweeps(b, b->gen, nullptr, first_sweep, s));
    const uint32_t zero = 0;
    HIP_TRY(hipMemcpyAsync(b->iter_dev, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (use_graph) {
        GraphGuard plain, with_reunit;
        if (reunit_every != 1) SUN_TRY(capture(s, plain, [&] { return batch_iteration(b, n_or, false, measure, s); }));
        if (reunit_every >= 1)
            SUN_TRY(capture(s, with_reunit, [&] { return batch_iteration(b, n_or, true, measure, s); }));
        for (int it = 0; it < n_iter; it++) {
            const bool ru = reunit_every > 0 && ((it + 1) % reunit_every) == 0;
            HIP_TRY(hipGraphLaunch(ru ? with_reunit.x : plain.x, s));
        }
    } else {
        for (int it = 0; it < n_iter; it++) {
            const bool ru = reunit_every > 0 && ((it + 1) % reunit_every) == 0;
            SUN_TRY(batch_iteration(b, n_or, ru, measure, s));
        }
    }
    // keep the host copy of the sweep counters in step with the device table
    for (int i = 0; i < n; i++) b->gen.h_update[i].sweep = first_sweep + (uint32_t)n_iter;
    if (measure) {
        std::vector<double2> h((size_t)n * n_iter);
        HIP_TRY(hipMemcpyAsync(h.data(), b->obs_dev, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (size_t k = 0; k < h.size(); k++) {
            obs[2 * k] = h[k].x;
            obs[2 * k + 1] = h[k].y;
        }
    } else {
        HIP_TRY(hipStreamSynchronize(s));
    }
    return SUN_OK;
}

extern "C" int sun_batch_generate(sun_batch* b, uint32_t first_sweep, int n_iter, int n_or, int reunit_every,
                                  int use_graph, void* stream, double* obs) {
    return guard_alloc("sun_batch_generate", [&] {
        return sun_batch_generate_body(b, first_sweep, n_iter, n_or, reunit_every, use_graph, stream, obs);
    });
}
