// sun_tune.hip -- launch-shape autotuner.  Replaces the reference's QUDA-derived
// tuner (tune.h:119-318, tune.cpp:158-414): for each (kernel, N, storage,
// precision, lattice extents, device arch) it times the real update kernel at every
// candidate block size on a scratch copy of the field (the reference's
// preTune/postTune Backup/Restore, monte.h:73-74), keeps the fastest, and persists
// the table in $SUN_RESOURCE_PATH/sun_tunecache.json (the reference's
// $CULQCD_RESOURCE_PATH/tunecache.csv).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <utility>
#include <vector>

#include "sun_internal.hpp"

namespace sun {

// candidate launch shapes, each with every LDS occupancy cap: linear mappings
// (b1 < 0) of `block` threads, and bricks of 2^b1 x 2^b2 x 2^b3 rows (x1, x2, x3)
// whose workgroup size follows from X0/2
struct Shape {
    int block, b1, b2, b3;
};
// The x3-deep bricks (b3 = 2..4 with thin x1, x2) can span the whole slab of x3 planes
// that one XCD's contiguous range of bricks covers (X3 / 8 planes), so that the x3
// neighbours of a brick are its own rows rather than a brick ~n1 n2 positions later.
static const Shape kShapes[] = {{64, -1, 0, 0},  {128, -1, 0, 0}, {256, -1, 0, 0}, {512, -1, 0, 0},
                                {0, 1, 1, 1},    {0, 2, 0, 2},    {0, 2, 1, 1},    {0, 1, 1, 2},
                                {0, 2, 2, 1},    {0, 2, 1, 2},    {0, 1, 2, 2},    {0, 2, 2, 2},
                                {0, 3, 1, 2},    {0, 2, 1, 3},    {0, 0, 0, 2},    {0, 0, 0, 3},
                                {0, 1, 0, 2},    {0, 0, 1, 2},    {0, 1, 0, 3},    {0, 0, 1, 3},
                                {0, 0, 0, 4}};
// dynamic LDS that no kernel reads: caps the workgroups per CU (160 KiB of LDS)
static const int kLds[] = {0, 16384, 32768, 40960, 54272, 65536, 81920};

std::string Tuner::key(const char* kernel, const sun_lattice_desc& d) const {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s|N%d|S%d|P%d|%dx%dx%dx%d|%s", kernel, d.nc, d.storage, d.precision, d.dims[0],
             d.dims[1], d.dims[2], d.dims[3], arch.c_str());
    return buf;
}

std::string Tuner::path() const {
    const char* dir = getenv("SUN_RESOURCE_PATH");
    if (!dir || !*dir) return std::string();
    return std::string(dir) + "/sun_tunecache.json";
}

// The cache file is a JSON object {"key": [block, b1, b2, b3, lds, piped, ms], ...},
// one entry per line.
void Tuner::load(int mode) {
    std::lock_guard<std::mutex> lk(mu_);
    mode_ = mode;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
    std::string p = path();
    if (p.empty()) return;
    std::ifstream in(p);
    std::string line;
    while (std::getline(in, line)) {
        size_t q0 = line.find('"');
        size_t q1 = q0 == std::string::npos ? q0 : line.find('"', q0 + 1);
        size_t b0 = line.find('[');
        if (q1 == std::string::npos || b0 == std::string::npos) continue;
        LaunchCfg c;
        float ms = 0.f;
        if (sscanf(line.c_str() + b0, "[%d, %d, %d, %d, %d, %d, %f]", &c.block, &c.b1, &c.b2, &c.b3, &c.lds, &c.piped,
                   &ms) != 7)
            continue;
        bool ok = false;
        for (const Shape& k : kShapes)
            ok |= k.b1 < 0 ? (c.b1 < 0 && k.block == c.block) : (k.b1 == c.b1 && k.b2 == c.b2 && k.b3 == c.b3);
        bool okl = false;
        for (int k : kLds) okl |= (k == c.lds);
        if (ok && okl) cache_[line.substr(q0 + 1, q1 - q0 - 1)] = {c, ms};
    }
}

void Tuner::save() {
    std::lock_guard<std::mutex> lk(mu_);
    std::string p = path();
    if (p.empty() || !dirty_) return;
    std::ofstream out(p);
    out << "{\n";
    size_t i = 0;
    for (auto& kv : cache_) {
        const LaunchCfg& c = kv.second.first;
        out << "  \"" << kv.first << "\": [" << c.block << ", " << c.b1 << ", " << c.b2 << ", " << c.b3 << ", "
            << c.lds << ", " << c.piped << ", " << kv.second.second << "]"
            << (++i < cache_.size() ? ",\n" : "\n");
    }
    out << "}\n";
    dirty_ = false;
}

LaunchCfg Tuner::lookup(const char* kernel, const sun_lattice_desc& d) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = cache_.find(key(kernel, d));
    if (it != cache_.end()) return it->second.first;
    return LaunchCfg();
}

bool Tuner::has(const char* kernel, const sun_lattice_desc& d) {
    std::lock_guard<std::mutex> lk(mu_);
    return cache_.count(key(kernel, d)) != 0;
}

void Tuner::record(const char* kernel, const sun_lattice_desc& d, const LaunchCfg& c, float ms) {
    std::lock_guard<std::mutex> lk(mu_);
    cache_[key(kernel, d)] = {c, ms};
    dirty_ = true;
}

// Times the four (parity 0, mu) launches of the heatbath or overrelaxation kernel
// for each candidate launch shape and stores the fastest in lat->cfg_hb / cfg_or.
int tune_lattice(sun_lattice* lat, hipStream_t s) {
    sun_ctx* ctx = lat->ctx;
    if (!ctx->tuner.enabled() || lat->tuned) return SUN_OK;
    void* backup = nullptr;
    HIP_TRY(hipMalloc(&backup, lat->bytes));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SUN_OK;
    hipError_t e = hipMemcpyAsync(backup, lat->buf, lat->bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const double bnc = lat->d.beta / lat->d.nc;
    // n launches over mu = 0..3 of candidate c, after one warm-up launch: average ms
    auto time_cand = [&](bool hb, const LaunchCfg& c, int n, float* ms) {
        e = lat->ops->update(lat->f, lat->g, 0, 0, hb, 0, nullptr, bnc, 1, c, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < n && e == hipSuccess; r++)
            e = lat->ops->update(lat->f, lat->g, 0, r & 3, hb, 0, nullptr, bnc, 1, c, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
        *ms /= (float)n;
        return e == hipSuccess;
    };
    LaunchCfg best_k[2];
    float best_k_ms[2] = {1e30f, 1e30f};
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        const bool hb = k == 0;
        // coarse pass: every (kernel, shape, LDS cap) candidate, 8 launches each; piped
        // 0 plain kernel, 1 pipelined by directions, 2 pipelined by half-directions
        std::vector<std::pair<LaunchCfg, float>> cands;
        for (int piped = 0; piped < 3 && e == hipSuccess; piped++) {
            if (piped && !lat->ops->has_piped) continue;
            const int max_block = piped ? lat->ops->piped_max_block : lat->ops->max_block;
            const int nshape = (int)(sizeof(kShapes) / sizeof(Shape)), nlds = (int)(sizeof(kLds) / sizeof(int));
            for (int cand = 0; cand < nshape * nlds; cand++) {
                const Shape& sh = kShapes[cand / nlds];
                LaunchCfg c;
                c.b1 = sh.b1;
                c.b2 = sh.b2;
                c.b3 = sh.b3;
                c.block = sh.b1 < 0 ? sh.block : brick_block(lat->g, sh.b1, sh.b2, sh.b3, max_block);
                c.lds = kLds[cand % nlds];
                c.piped = piped;
                if (c.block == 0 || c.block > max_block) continue;
                float ms = 0.f;
                if (!time_cand(hb, c, 8, &ms)) break;
                if (ctx->verbosity >= SUN_VERBOSE)
                    fprintf(stderr, "[sun tune] %s piped=%d block=%d brick=%d,%d,%d lds=%d %.4f ms\n",
                            hb ? "heatbath" : "overrelax", piped, c.block, c.b1, c.b2, c.b3, c.lds, ms);
                cands.push_back({c, ms});
            }
        }
        if (e != hipSuccess || cands.empty()) break;
        // refinement: the candidates within 4% of the fastest (at most 8, in candidate
        // order) timed again in three interleaved rounds of 16 launches; the lowest
        // total wins, a later candidate only by more than 0.5%.  Near-equal shapes
        // otherwise swap places from run to run with the timing noise of 8 launches.
        float coarse_best = 1e30f;
        for (const auto& c : cands) coarse_best = std::min(coarse_best, c.second);
        std::vector<LaunchCfg> fin;
        for (const auto& c : cands)
            if (c.second <= coarse_best * 1.04f && fin.size() < 8) fin.push_back(c.first);
        std::vector<float> total(fin.size(), 0.f);
        for (int round = 0; round < 3 && e == hipSuccess; round++)
            for (size_t i = 0; i < fin.size() && e == hipSuccess; i++) {
                float ms = 0.f;
                if (time_cand(hb, fin[i], 16, &ms)) total[i] += ms / 3.f;
            }
        if (e != hipSuccess) break;
        size_t win = 0;
        for (size_t i = 1; i < fin.size(); i++)
            if (total[i] < total[win] * 0.995f) win = i;
        best_k[k] = fin[win];
        best_k_ms[k] = total[win];
    }
    if (e == hipSuccess) {
        lat->cfg_hb = best_k[0];
        lat->cfg_or = best_k[1];
        ctx->tuner.record("heatbath", lat->d, best_k[0], best_k_ms[0]);
        ctx->tuner.record("overrelax", lat->d, best_k[1], best_k_ms[1]);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(lat->buf, backup, lat->bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(SUN_ERR_HIP, "autotuner: %s", hipGetErrorString(e));
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(backup);
    if (rc == SUN_OK) lat->tuned = true;
    return rc;
}

}  // namespace sun
