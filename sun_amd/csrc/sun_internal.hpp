// sun_internal.hpp -- handle structures and helpers shared by the C-ABI sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "sun.h"
#include "sun_ops.hpp"

namespace sun {

int fail(int code, const char* fmt, ...);

// Runs the body of an entry point whose host buffers scale with the lattice or with
// a caller's count: a failed host allocation becomes SUN_ERR_NOMEM instead of an
// exception crossing the C ABI.
template <class F>
int guard_alloc(const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(SUN_ERR_NOMEM, "%s: host allocation failed", what);
    }
}

#define HIP_TRY(x)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            hipGetLastError();                                                                     \
            return ::sun::fail(SUN_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,                \
                               hipGetErrorString(e_));                                             \
        }                                                                                          \
    } while (0)

#define HIP_TRY_CLEANUP(x, cleanup)                                                                \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            hipGetLastError();                                                                     \
            cleanup;                                                                               \
            return ::sun::fail(SUN_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,                \
                               hipGetErrorString(e_));                                             \
        }                                                                                          \
    } while (0)

#define SUN_TRY(x)              \
    do {                        \
        int rc_ = (x);          \
        if (rc_ != SUN_OK) return rc_; \
    } while (0)

// Launch-shape autotuner (replaces the reference's tune.cpp:158-414): per
// (kernel, N, storage, precision, volume, device) the fastest block size, found by
// timing the real kernel on a scratch copy of the field, cached in memory and in
// $SUN_RESOURCE_PATH/sun_tunecache.json.
class Tuner {
   public:
    void load(int mode);
    void save();
    LaunchCfg lookup(const char* kernel, const sun_lattice_desc& d);
    bool has(const char* kernel, const sun_lattice_desc& d);
    void record(const char* kernel, const sun_lattice_desc& d, const LaunchCfg& c, float ms);
    bool enabled() const { return mode_ != 0; }
    std::string arch;

   private:
    std::string key(const char* kernel, const sun_lattice_desc& d) const;
    std::string path() const;
    int mode_ = 0;
    bool dirty_ = false;
    std::map<std::string, std::pair<LaunchCfg, float>> cache_;
    std::mutex mu_;
};

}  // namespace sun

struct sun_ctx {
    int device = 0;
    int verbosity = 0;
    int tune = 0;
    std::string name, gcn_arch;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    sun::Tuner tuner;
    // this device's copy of the allocation-free entry points' scratch array (its
    // device address, looked up once per device at sun_init)
    void* scratch = nullptr;
};

namespace sun {
// size of the device scratch array of the allocation-free entry points (one per
// device, held until the process ends): sun_wilson_loops_noalloc keeps its Wilson-line
// field (3V links in 18-real storage) and its partial sums there
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
// device result slots of a lattice: plaquette and Polyakov of sun_generate, or the
// three of the link monitor
constexpr int kResSlots = 3;
}  // namespace sun

struct sun_lattice {
    sun_ctx* ctx = nullptr;
    sun_lattice_desc d{};
    const sun::OpsTable* ops = nullptr;
    sun::Geom g{};
    sun::Field f{};
    void* buf = nullptr;
    size_t bytes = 0;
    double2* partials = nullptr;   // per-block reduction scratch
    int npart = 0;
    double2* dres = nullptr;       // device results (kResSlots)
    double2* hres = nullptr;       // pinned host copy (kResSlots), null for server lattices
    uint32_t* sweep_dev = nullptr; // device sweep counter for graph replay
    double2* obs_dev = nullptr;    // device observable history of sun_generate
    int obs_cap = 0;
    sun::LaunchCfg cfg_hb, cfg_or;
    bool tuned = false;
    bool external = false;         // memory belongs to a server arena: never hipFree'd
    double2* hist = nullptr;       // observable history (server jobs), in the arena
};

// Link bytes of one update chunk of a job-server group: the lattices of a chunk stay
// in the GPU's 256 MB memory-side cache (MI355X Infinity Cache) through the chunk's
// consecutive launches, so the links a launch reads again (neighbours shared across
// workgroups and XCDs, the same field in the next (parity, mu) launch) come from there
// instead of HBM.  24^4 FP32-12 lattices, one overrelaxation sweep: 4 in one batch
// (256 MB) 0.62 of 8 TB/s, 18 (1.15 GB) 0.55.  The 64-lattice benchmark mix (1.36 GB)
// per iteration, chunks run one after another, by chunk budget: none 9.7-9.8e9 link
// updates/s, 96 MB 8.5e9, 128 MB 9.0e9, 192 MB 9.9-10.1e9, 256 MB 10.3-10.6e9, 320 MB
// 10.1-10.3e9, 384 MB 9.6-9.8e9 (smaller chunks: launches too small to fill the GPU).
constexpr size_t kCacheChunkBytes = size_t(256) << 20;

// sun_batch runs its chunks on kBatchBranches streams at once (chunk i on branch
// i % kBatchBranches), so one chunk's launches fill the tails of another's, with
// chunks of a third of the cache: the chunks in flight together still fit it.  The
// 64-lattice mix per iteration (ms, one box): one branch of 256 MB chunks 15.3; two
// branches of 96 / 128 / 160 MB 15.1 / 14.9 / 15.7, of 256 MB 17.3; three of 72 / 85 /
// 100 MB 14.5-14.7; four of 56 / 64 MB 15.0-15.3; five of 51 MB 16.6-16.8.
constexpr int kBatchBranches = 3;
constexpr size_t kBatchChunkBytes = kCacheChunkBytes / kBatchBranches;

struct sun_batch {
    sun_ctx* ctx = nullptr;
    std::vector<sun_lattice*> lats;
    const sun::OpsTable* ops = nullptr;
    // chunks: consecutive lattices whose link fields together stay within a budget (one
    // larger lattice is a chunk of its own), run back to back while their fields sit in
    // the memory-side cache.  blocks_* are the chunk's update (half/256), reunit
    // (4V/256) and plaquette (V/256) blocks.
    struct Chunk {
        int first = 0, n = 0, blocks = 0, blocks_reunit = 0, blocks_plaq = 0;
    };
    // one chunking of the batch: its chunks, the device tables of its lattices (whose
    // first_block counts from the start of the lattice's chunk), the host copy of the
    // update table, and the number of streams its chunks run on
    struct Schedule {
        std::vector<Chunk> chunks;
        sun::BatchEntry* t_update = nullptr;
        sun::BatchEntry* t_reunit = nullptr;
        sun::BatchEntry* t_plaq = nullptr;
        std::vector<sun::BatchEntry> h_update;
        int branches = 1;
    };
    // sun_batch_generate: chunks of kBatchChunkBytes on kBatchBranches streams (an
    // iteration's heatbath, overrelaxation, reunitarisation and plaquette launches of
    // different chunks overlap); the single-sweep entry points: chunks of
    // kCacheChunkBytes one after another (one overrelaxation sweep alone ran at 0.65 of
    // 8 TB/s that way and 0.60 on the branches of generate's chunks)
    Schedule gen, step;
    int piped[2] = {0, 0};  // staple kernel of the heatbath / overrelaxation launches
    int brick[2] = {0, 0};  // their lattices' update bricks instead of linear order
    int lds[2] = {0, 0};    // unused dynamic LDS per workgroup (occupancy cap)
    std::vector<sun::BatchEntry> h_update;
    double2* partials = nullptr;
    double2* d_out = nullptr;
    double2* h_out = nullptr;
    double2* obs_dev = nullptr;
    int obs_cap = 0;
    uint32_t* iter_dev = nullptr;
    // branches 1.. of the chunk schedule (branch 0 is the caller's stream) and the
    // events that fork them from it and join them back
    hipStream_t branch[kBatchBranches] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kBatchBranches] = {};
};

namespace sun {
int check_desc(const sun_lattice_desc* d);
size_t link_bytes(const sun_lattice_desc* d);
// plane view of nrec link records of lb bytes at base (the layout of a lattice's
// field: 16-byte records per plane, an 8-byte tail plane when lb % 16 == 8)
Field make_field(void* base, size_t nrec, size_t lb);
int lattice_alloc(sun_ctx* ctx, const sun_lattice_desc* d, sun_lattice** out);
// bytes of a lattice placed in caller memory: link field, reduction scratch,
// result slots, sweep counter and an n_hist-row observable history
size_t lattice_footprint(const sun_lattice_desc* d, int n_hist);
// a lattice whose every device buffer is carved from mem (lattice_footprint bytes,
// 256-byte aligned); lattice_free then only releases the host structure
int lattice_in_memory(sun_ctx* ctx, const sun_lattice_desc* d, char* mem, int n_hist, sun_lattice** out);
void lattice_free(sun_lattice* L);
hipStream_t pick_stream(sun_lattice* lat, void* stream);
int heatbath_sweep(sun_lattice* lat, uint64_t seed, uint32_t sweep, const uint32_t* sweep_dev, hipStream_t s);
int overrelax_sweep(sun_lattice* lat, hipStream_t s);
int plaquette_to(sun_lattice* lat, hipStream_t s, double2* dev_out);
int polyakov_to(sun_lattice* lat, hipStream_t s, double2* dev_out);
int link_monitor_to(sun_lattice* lat, hipStream_t s, double2* dev_out);
int tune_lattice(sun_lattice* lat, hipStream_t s);
// small helper kernels (sun_common.hip)
hipError_t launch_advance_counter(uint32_t* ctr, hipStream_t s);
hipError_t launch_record_obs(const double2* src, int n, double2* history, const uint32_t* ctr, uint32_t first,
                             hipStream_t s);
hipError_t launch_batch_advance(BatchEntry* table, int n, hipStream_t s);
// native checkpoint of a lattice on an open file (sun_io.hip)
int write_native(FILE* f, sun_lattice* L, uint64_t seed, uint32_t next_sweep, const char* path);
int read_native(FILE* f, const sun_lattice_desc& d, size_t bytes, std::vector<unsigned char>& payload,
                uint64_t* seed, uint32_t* next_sweep, const char* path);
bool write_native_bytes(FILE* f, const sun_job_desc& d, const std::vector<unsigned char>& payload, uint64_t seed,
                        uint32_t next_sweep);

// Observables a generation run may report: finite, plaquette components and
// |Polyakov loop| at most 1 (Re Tr U / N of a unitary U; the slack covers FP32 links
// between reunitarisations).  A field that has overflowed to NaN or drifted off
// SU(N) fails this.
constexpr double kObsSlack = 1e-5;
inline bool observables_in_range(double2 plaq, double2 poly) {
    const double lim = 1.0 + kObsSlack;
    return std::isfinite(plaq.x) && std::isfinite(plaq.y) && std::isfinite(poly.x) && std::isfinite(poly.y) &&
           std::fabs(plaq.x) <= lim && std::fabs(plaq.y) <= lim && std::hypot(poly.x, poly.y) <= lim;
}
}  // namespace sun
