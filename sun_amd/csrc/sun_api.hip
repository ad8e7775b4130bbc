// sun_api.hip -- the C ABI declared in include/sun.h: contexts, lattice handles,
// error reporting, dispatch to the per-variant kernel tables, measurement, the
// generation driver (optionally one hipGraph per iteration) and batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sun.h"
#include "sun_internal.hpp"

using namespace sun;

// ---------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;
}

int sun::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" const char* sun_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* sun_version(void) { return "sun-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------
// Scratch of the allocation-free entry points (sun_wilson_loops_noalloc): a device
// array of fixed size instead of hipMalloc, for callers that cannot allocate (code
// under stream capture, serving loops).  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU is looked up once, at
// the first sun_init on that device, and only that address is used afterwards.
static __device__ __align__(256) unsigned char scratch[sun::kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// the device address of this device's scratch array; the device is current
int lookup_scratch(int device, void** ptr) {
    if (device >= kMaxDevices) return fail(SUN_ERR_UNSUPPORTED, "sun_init: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SUN_ERR_HIP, "sun_init: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_dev[device] = p;
    }
    *ptr = g_scratch_dev[device];
    return SUN_OK;
}
}  // namespace

extern "C" int sun_init(int device, int verbosity, int tune, sun_ctx** out) {
    if (!out) return fail(SUN_ERR_INVALID, "sun_init: out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SUN_ERR_INVALID, "sun_init: device %d of %d", device, ndev);
    HIP_TRY(hipSetDevice(device));
    sun_ctx* c = new (std::nothrow) sun_ctx();
    if (!c) return fail(SUN_ERR_NOMEM, "sun_init: host allocation");
    c->device = device;
    c->verbosity = verbosity;
    c->tune = tune;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->name = prop.name;
        c->gcn_arch = prop.gcnArchName;
        c->num_cu = prop.multiProcessorCount;
    }
    const int rc = lookup_scratch(device, &c->scratch);
    if (rc != SUN_OK) {
        delete c;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(SUN_ERR_HIP, "sun_init: hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->tuner.load(tune);
    if (verbosity >= SUN_SUMMARIZE)
        fprintf(stderr, "[sun] device %d: %s (%s, %d CUs)\n", device, c->name.c_str(), c->gcn_arch.c_str(), c->num_cu);
    *out = c;
    return SUN_OK;
}

extern "C" int sun_finalize(sun_ctx* ctx) {
    if (!ctx) return fail(SUN_ERR_INVALID, "sun_finalize: NULL context");
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->tuner.save();
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SUN_OK;
}

extern "C" int sun_device_name(sun_ctx* ctx, char* buf, int len) {
    if (!ctx || !buf || len <= 0) return fail(SUN_ERR_INVALID, "sun_device_name: bad argument");
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", ctx->name.c_str(), ctx->gcn_arch.c_str(), ctx->num_cu);
    return SUN_OK;
}

// ---------------------------------------------------------------------------------
// lattices
// ---------------------------------------------------------------------------------
int sun::check_desc(const sun_lattice_desc* d) {
    if (!d) return fail(SUN_ERR_INVALID, "lattice descriptor is NULL");
    if (d->nc < 2 || d->nc > 4) return fail(SUN_ERR_UNSUPPORTED, "nc=%d: supported are 2, 3, 4", d->nc);
    long long V = 1;
    for (int i = 0; i < 4; i++) {
        if (d->dims[i] < 2 || (d->dims[i] & 1))
            return fail(SUN_ERR_INVALID, "dims[%d]=%d: extents must be even and >= 2", i, d->dims[i]);
        V *= d->dims[i];
        // a plane holds 4V records of 16 bytes addressed by 32-bit byte offsets; checked
        // after every factor, so the product (at most 2^26 * 2^31) cannot overflow
        if (4 * V * 16 > 0xFFFFFFFFll)
            return fail(SUN_ERR_INVALID, "volume too large at dims[%d]=%d (4V x 16 bytes must fit in 32 bits)", i,
                        d->dims[i]);
    }
    if (d->storage != SUN_STORE_18 && d->storage != SUN_STORE_12 && d->storage != SUN_STORE_8)
        return fail(SUN_ERR_INVALID, "storage=%d: use 18, 12 or 8", d->storage);
    if (d->nc != 3 && d->storage != SUN_STORE_18)
        return fail(SUN_ERR_UNSUPPORTED, "12/8-real storage is SU(3) only (gaugearray.h:127-137)");
    if (d->precision != SUN_PREC_F32 && d->precision != SUN_PREC_F64)
        return fail(SUN_ERR_INVALID, "precision=%d: use 4 (f32) or 8 (f64)", d->precision);
    if (!(d->beta > 0.0)) return fail(SUN_ERR_INVALID, "beta=%g must be positive", d->beta);
    return SUN_OK;
}

size_t sun::link_bytes(const sun_lattice_desc* d) {
    const int nreal = d->storage == SUN_STORE_18 ? 2 * d->nc * d->nc : d->storage;
    return (size_t)nreal * (size_t)d->precision;
}

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

Field sun::make_field(void* base, size_t nrec, size_t lb) {
    Field f;
    memset(&f, 0, sizeof(f));
    const int nfull = (int)(lb / 16);
    const int tail = (int)(lb % 16);
    for (int c = 0; c < nfull; c++) f.p[c] = (char*)base + (size_t)c * nrec * 16;
    if (tail) f.p[nfull] = (char*)base + (size_t)nfull * nrec * 16;
    return f;
}

static void set_planes(sun_lattice* L) { L->f = make_field(L->buf, 4 * (size_t)L->g.V, link_bytes(&L->d)); }

extern "C" size_t sun_lattice_bytes(const sun_lattice_desc* d) {
    if (check_desc(d) != SUN_OK) return 0;
    size_t V = (size_t)d->dims[0] * d->dims[1] * d->dims[2] * d->dims[3];
    return 4 * V * link_bytes(d);
}

int sun::lattice_alloc(sun_ctx* ctx, const sun_lattice_desc* d, sun_lattice** out) {
    int rc = check_desc(d);
    if (rc) return rc;
    const OpsTable* ops = find_ops(d->nc, d->storage, d->precision);
    if (!ops) return fail(SUN_ERR_UNSUPPORTED, "no kernels for nc=%d storage=%d precision=%d", d->nc, d->storage,
                          d->precision);
    sun_lattice* L = new (std::nothrow) sun_lattice();
    if (!L) return fail(SUN_ERR_NOMEM, "host allocation");
    L->ctx = ctx;
    L->d = *d;
    L->ops = ops;
    L->g = make_geom(d->dims);
    const size_t nl = 4 * (size_t)L->g.V;
    const size_t lb = link_bytes(d);
    const int nfull = (int)(lb / 16);
    const int tail = (int)(lb % 16);
    L->bytes = nl * lb;
    HIP_TRY_CLEANUP(hipSetDevice(ctx->device), delete L);
    hipError_t e = hipMalloc(&L->buf, L->bytes);
    if (e != hipSuccess) {
        delete L;
        hipGetLastError();
        return fail(SUN_ERR_NOMEM, "hipMalloc(%zu bytes) for links: %s", nl * lb, hipGetErrorString(e));
    }
    (void)nfull;
    (void)tail;
    set_planes(L);
    L->npart = partial_slots(L->g);
    e = hipMalloc(&L->partials, sizeof(double2) * (size_t)L->npart);
    if (e == hipSuccess) e = hipMalloc(&L->dres, sizeof(double2) * kResSlots);
    if (e == hipSuccess) e = hipMalloc(&L->sweep_dev, sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(&L->hres, sizeof(double2) * kResSlots, hipHostMallocDefault);
    if (e != hipSuccess) {
        hipGetLastError();
        lattice_free(L);
        return fail(SUN_ERR_NOMEM, "scratch allocation: %s", hipGetErrorString(e));
    }
    e = hipMemsetAsync(L->buf, 0, L->bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        lattice_free(L);
        return fail(SUN_ERR_HIP, "clearing link field: %s", hipGetErrorString(e));
    }
    L->cfg_hb = ctx->tuner.lookup("heatbath", *d);
    L->cfg_or = ctx->tuner.lookup("overrelax", *d);
    L->tuned = ctx->tuner.has("heatbath", *d) && ctx->tuner.has("overrelax", *d);
    *out = L;
    return SUN_OK;
}

size_t sun::lattice_footprint(const sun_lattice_desc* d, int n_hist) {
    if (check_desc(d) != SUN_OK) return 0;
    const Geom g = make_geom(d->dims);
    const size_t field = 4 * (size_t)g.V * link_bytes(d);
    const size_t part = sizeof(double2) * (size_t)partial_slots(g);
    return align256(field) + align256(part) + align256(kResSlots * sizeof(double2)) + align256(sizeof(uint32_t)) +
           align256(sizeof(double2) * 2 * (size_t)(n_hist > 0 ? n_hist : 1));
}

int sun::lattice_in_memory(sun_ctx* ctx, const sun_lattice_desc* d, char* mem, int n_hist, sun_lattice** out) {
    int rc = check_desc(d);
    if (rc) return rc;
    const OpsTable* ops = find_ops(d->nc, d->storage, d->precision);
    if (!ops) return fail(SUN_ERR_UNSUPPORTED, "no kernels for nc=%d storage=%d precision=%d", d->nc, d->storage,
                          d->precision);
    sun_lattice* L = new (std::nothrow) sun_lattice();
    if (!L) return fail(SUN_ERR_NOMEM, "host allocation");
    L->ctx = ctx;
    L->d = *d;
    L->ops = ops;
    L->g = make_geom(d->dims);
    L->external = true;
    const size_t field = 4 * (size_t)L->g.V * link_bytes(d);
    L->bytes = field;
    L->npart = partial_slots(L->g);
    char* p = mem;
    L->buf = p;
    p += align256(field);
    L->partials = (double2*)p;
    p += align256(sizeof(double2) * (size_t)L->npart);
    L->dres = (double2*)p;
    p += align256(kResSlots * sizeof(double2));
    L->sweep_dev = (uint32_t*)p;
    p += align256(sizeof(uint32_t));
    L->hist = (double2*)p;
    set_planes(L);
    L->cfg_hb = ctx->tuner.lookup("heatbath", *d);
    L->cfg_or = ctx->tuner.lookup("overrelax", *d);
    L->tuned = true;  // server lattices never tune: tuning needs a scratch copy outside the budget
    *out = L;
    return SUN_OK;
}

void sun::lattice_free(sun_lattice* L) {
    if (!L) return;
    if (L->external) {
        delete L;
        return;
    }
    if (L->ctx) hipSetDevice(L->ctx->device);
    if (L->buf) hipFree(L->buf);
    if (L->partials) hipFree(L->partials);
    if (L->dres) hipFree(L->dres);
    if (L->sweep_dev) hipFree(L->sweep_dev);
    if (L->obs_dev) hipFree(L->obs_dev);
    if (L->hres) hipHostFree(L->hres);
    delete L;
}

extern "C" int sun_lattice_create(sun_ctx* ctx, const sun_lattice_desc* d, sun_lattice** out) {
    if (!ctx || !out) return fail(SUN_ERR_INVALID, "sun_lattice_create: NULL argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return lattice_alloc(ctx, d, out);
}

extern "C" int sun_lattice_destroy(sun_lattice* lat) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_lattice_destroy: NULL lattice");
    hipSetDevice(lat->ctx->device);
    hipDeviceSynchronize();
    lattice_free(lat);
    return SUN_OK;
}

extern "C" int sun_lattice_copy(sun_lattice* dst, const sun_lattice* src, void* stream) {
    if (!dst || !src) return fail(SUN_ERR_INVALID, "sun_lattice_copy: NULL lattice");
    if (dst->ops != src->ops || memcmp(dst->d.dims, src->d.dims, sizeof(dst->d.dims)) != 0)
        return fail(SUN_ERR_INVALID, "sun_lattice_copy: lattices differ in N, extents, storage or precision");
    HIP_TRY(hipMemcpyAsync(dst->buf, src->buf, src->bytes, hipMemcpyDeviceToDevice, pick_stream(dst, stream)));
    return SUN_OK;
}

extern "C" int sun_lattice_convert(sun_lattice* dst, const sun_lattice* src, void* stream) {
    if (!dst || !src) return fail(SUN_ERR_INVALID, "sun_lattice_convert: NULL lattice");
    if (dst->d.nc != src->d.nc || memcmp(dst->d.dims, src->d.dims, sizeof(dst->d.dims)) != 0)
        return fail(SUN_ERR_INVALID, "sun_lattice_convert: lattices differ in N or extents");
    if (dst->ctx->device != src->ctx->device)
        return fail(SUN_ERR_INVALID, "sun_lattice_convert: lattices on different devices");
    hipStream_t s = pick_stream(dst, stream);
    // every link through the canonical FP64 form: the source's stored rows decoded
    // (12/8-real reconstructed) and widened, then encoded in the destination's storage
    // and precision -- the same two kernels as a host round trip, without the host
    const size_t n = 4 * (size_t)src->g.V * src->d.nc * src->d.nc * 2;
    double* tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, n * sizeof(double)));
    int rc = SUN_OK;
    hipError_t e = src->ops->to_canonical(src->f, 4 * src->g.V, tmp, s);
    if (e == hipSuccess) e = dst->ops->from_canonical(dst->f, 4 * dst->g.V, tmp, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(SUN_ERR_HIP, "sun_lattice_convert: %s", hipGetErrorString(e));
    hipFree(tmp);
    return rc;
}

extern "C" int sun_lattice_get_launch(const sun_lattice* lat, int kernel, sun_launch* out) {
    if (!lat || !out || (kernel != 0 && kernel != 1)) return fail(SUN_ERR_INVALID, "sun_lattice_get_launch: bad argument");
    const LaunchCfg& c = kernel == 0 ? lat->cfg_hb : lat->cfg_or;
    out->block = c.b1 >= 0 ? brick_block(lat->g, c.b1, c.b2, c.b3, c.piped ? lat->ops->piped_max_block
                                                                              : lat->ops->max_block)
                           : c.block;
    out->brick[0] = c.b1;
    out->brick[1] = c.b2;
    out->brick[2] = c.b3;
    out->lds = c.lds;
    out->piped = c.piped;
    return SUN_OK;
}

// a caller's launch shape checked against the kernels of `ops` on geometry g
static int launch_from_user(const OpsTable* ops, const Geom& g, const int dims[4], const sun_launch* cfg,
                            const char* what, LaunchCfg* out) {
    if (cfg->piped && !ops->has_piped)
        return fail(SUN_ERR_INVALID, "%s: no pipelined kernel for this N/storage/precision", what);
    const int max_block = cfg->piped ? ops->piped_max_block : ops->max_block;
    if (cfg->piped < 0 || cfg->piped > 2) return fail(SUN_ERR_INVALID, "%s: piped %d (0, 1 or 2)", what, cfg->piped);
    LaunchCfg c;
    c.piped = cfg->piped;
    c.lds = cfg->lds;
    if (cfg->lds < 0 || cfg->lds > 65536 * 2) return fail(SUN_ERR_INVALID, "%s: lds %d", what, cfg->lds);
    if (cfg->brick[0] >= 0) {
        const int block = brick_block(g, cfg->brick[0], cfg->brick[1], cfg->brick[2], max_block);
        if (!block)
            return fail(SUN_ERR_INVALID, "%s: brick %d,%d,%d does not tile %dx%dx%dx%d", what, cfg->brick[0],
                        cfg->brick[1], cfg->brick[2], dims[0], dims[1], dims[2], dims[3]);
        c.b1 = cfg->brick[0];
        c.b2 = cfg->brick[1];
        c.b3 = cfg->brick[2];
        c.block = block;
    } else {
        if (cfg->block < 64 || cfg->block > max_block || (cfg->block & 63))
            return fail(SUN_ERR_INVALID, "%s: block %d (multiple of 64, <= %d)", what, cfg->block, max_block);
        c.block = cfg->block;
    }
    *out = c;
    return SUN_OK;
}

extern "C" int sun_lattice_set_launch(sun_lattice* lat, int kernel, const sun_launch* cfg) {
    if (!lat || !cfg || (kernel != 0 && kernel != 1)) return fail(SUN_ERR_INVALID, "sun_lattice_set_launch: bad argument");
    LaunchCfg c;
    SUN_TRY(launch_from_user(lat->ops, lat->g, lat->d.dims, cfg, "sun_lattice_set_launch", &c));
    (kernel == 0 ? lat->cfg_hb : lat->cfg_or) = c;
    lat->tuned = true;
    return SUN_OK;
}

extern "C" int sun_ctx_set_launch(sun_ctx* ctx, const sun_lattice_desc* desc, int kernel, const sun_launch* cfg) {
    if (!ctx || !desc || !cfg || (kernel != 0 && kernel != 1))
        return fail(SUN_ERR_INVALID, "sun_ctx_set_launch: bad argument");
    SUN_TRY(check_desc(desc));
    const OpsTable* ops = find_ops(desc->nc, desc->storage, desc->precision);
    if (!ops)
        return fail(SUN_ERR_UNSUPPORTED, "no kernels for nc=%d storage=%d precision=%d", desc->nc, desc->storage,
                    desc->precision);
    LaunchCfg c;
    SUN_TRY(launch_from_user(ops, make_geom(desc->dims), desc->dims, cfg, "sun_ctx_set_launch", &c));
    ctx->tuner.record(kernel == 0 ? "heatbath" : "overrelax", *desc, c, 0.f);
    return SUN_OK;
}

hipStream_t sun::pick_stream(sun_lattice* lat, void* stream) {
    return stream ? (hipStream_t)stream : lat->ctx->stream;
}

extern "C" int sun_cold_start(sun_lattice* lat, void* stream) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_cold_start: NULL lattice");
    if (lat->d.storage == SUN_STORE_8)
        return fail(SUN_ERR_UNSUPPORTED,
                    "8-real storage cannot hold a cold start: reconstruction is singular at |e00|=1 "
                    "(gaugearray.cu:293-297)");
    HIP_TRY(lat->ops->cold_start(lat->f, 4 * lat->g.V, pick_stream(lat, stream)));
    return SUN_OK;
}

extern "C" int sun_hot_start(sun_lattice* lat, uint64_t seed, void* stream) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_hot_start: NULL lattice");
    HIP_TRY(lat->ops->hot_start(lat->f, lat->g, seed, pick_stream(lat, stream)));
    return SUN_OK;
}

// ---------------------------------------------------------------------------------
// updates
// ---------------------------------------------------------------------------------
extern "C" int sun_heatbath_sub(sun_lattice* lat, uint64_t seed, uint32_t sweep, int parity, int mu, void* stream) {
    if (!lat || parity < 0 || parity > 1 || mu < 0 || mu > 3)
        return fail(SUN_ERR_INVALID, "sun_heatbath_sub: bad argument");
    HIP_TRY(lat->ops->update(lat->f, lat->g, parity, mu, true, sweep, nullptr, lat->d.beta / lat->d.nc, seed,
                             lat->cfg_hb, pick_stream(lat, stream)));
    return SUN_OK;
}

extern "C" int sun_overrelax_sub(sun_lattice* lat, int parity, int mu, void* stream) {
    if (!lat || parity < 0 || parity > 1 || mu < 0 || mu > 3)
        return fail(SUN_ERR_INVALID, "sun_overrelax_sub: bad argument");
    HIP_TRY(lat->ops->update(lat->f, lat->g, parity, mu, false, 0, nullptr, 0.0, 0, lat->cfg_or,
                             pick_stream(lat, stream)));
    return SUN_OK;
}

int sun::heatbath_sweep(sun_lattice* lat, uint64_t seed, uint32_t sweep, const uint32_t* sweep_dev, hipStream_t s) {
    const double bnc = lat->d.beta / lat->d.nc;  // constants.cu:443
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++)
            HIP_TRY(lat->ops->update(lat->f, lat->g, parity, mu, true, sweep, sweep_dev, bnc, seed, lat->cfg_hb, s));
    return SUN_OK;
}

int sun::overrelax_sweep(sun_lattice* lat, hipStream_t s) {
    for (int parity = 0; parity < 2; parity++)
        for (int mu = 0; mu < 4; mu++)
            HIP_TRY(lat->ops->update(lat->f, lat->g, parity, mu, false, 0, nullptr, 0.0, 0, lat->cfg_or, s));
    return SUN_OK;
}

extern "C" int sun_heatbath(sun_lattice* lat, uint64_t seed, uint32_t sweep, void* stream) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_heatbath: NULL lattice");
    hipStream_t s = pick_stream(lat, stream);
    SUN_TRY(tune_lattice(lat, s));
    return heatbath_sweep(lat, seed, sweep, nullptr, s);
}

extern "C" int sun_overrelax(sun_lattice* lat, void* stream) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_overrelax: NULL lattice");
    hipStream_t s = pick_stream(lat, stream);
    SUN_TRY(tune_lattice(lat, s));
    return overrelax_sweep(lat, s);
}

extern "C" int sun_reunitarize(sun_lattice* lat, void* stream) {
    if (!lat) return fail(SUN_ERR_INVALID, "sun_reunitarize: NULL lattice");
    HIP_TRY(lat->ops->reunitarize(lat->f, 4 * lat->g.V, pick_stream(lat, stream)));
    return SUN_OK;
}

// ---------------------------------------------------------------------------------
// observables
// ---------------------------------------------------------------------------------
int sun::plaquette_to(sun_lattice* lat, hipStream_t s, double2* dev_out) {
    const double norm = 1.0 / ((double)(3 * lat->d.nc) * (double)lat->g.V);  // plaquette_cub.cu:234
    HIP_TRY(lat->ops->plaquette_partial(lat->f, lat->g, lat->partials, s));
    HIP_TRY(launch_finalize(lat->partials, plaquette_blocks(lat->g), norm, norm, dev_out, s));
    return SUN_OK;
}

int sun::polyakov_to(sun_lattice* lat, hipStream_t s, double2* dev_out) {
    const double ts = (double)lat->g.X[0] * lat->g.X[1] * lat->g.X[2];
    const double norm = 1.0 / ((double)lat->d.nc * ts);  // polyakovloop.cu:651
    HIP_TRY(lat->ops->polyakov_partial(lat->f, lat->g, lat->partials, s));
    HIP_TRY(launch_finalize(lat->partials, polyakov_blocks(lat->g), norm, norm, dev_out, s));
    return SUN_OK;
}

static int read_result(sun_lattice* lat, hipStream_t s, double out[2]) {
    HIP_TRY(hipMemcpyAsync(lat->hres, lat->dres, sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out[0] = lat->hres[0].x;
    out[1] = lat->hres[0].y;
    return SUN_OK;
}

extern "C" int sun_plaquette(sun_lattice* lat, void* stream, double out[2]) {
    if (!lat || !out) return fail(SUN_ERR_INVALID, "sun_plaquette: NULL argument");
    hipStream_t s = pick_stream(lat, stream);
    int rc = plaquette_to(lat, s, lat->dres);
    return rc ? rc : read_result(lat, s, out);
}

extern "C" int sun_polyakov(sun_lattice* lat, void* stream, double out[2]) {
    if (!lat || !out) return fail(SUN_ERR_INVALID, "sun_polyakov: NULL argument");
    hipStream_t s = pick_stream(lat, stream);
    int rc = polyakov_to(lat, s, lat->dres);
    return rc ? rc : read_result(lat, s, out);
}

extern "C" int sun_plaquette_async(sun_lattice* lat, void* stream, double* dev_out) {
    if (!lat || !dev_out) return fail(SUN_ERR_INVALID, "sun_plaquette_async: NULL argument");
    return plaquette_to(lat, pick_stream(lat, stream), (double2*)dev_out);
}

extern "C" int sun_polyakov_async(sun_lattice* lat, void* stream, double* dev_out) {
    if (!lat || !dev_out) return fail(SUN_ERR_INVALID, "sun_polyakov_async: NULL argument");
    return polyakov_to(lat, pick_stream(lat, stream), (double2*)dev_out);
}

// link monitors: trace and determinant means, unitarity deviations
int sun::link_monitor_to(sun_lattice* lat, hipStream_t s, double2* dev_out) {
    const double V = (double)lat->g.V;
    HIP_TRY(lat->ops->monitor_partial(lat->f, lat->g, lat->partials, s));
    // linktrsum.cu:137 (4 N V), linkdetsum.cu:93 (4 V)
    HIP_TRY(launch_monitor_finalize(lat->partials, monitor_blocks(lat->g), 1.0 / (4.0 * lat->d.nc * V),
                                    1.0 / (4.0 * V), dev_out, s));
    return SUN_OK;
}

extern "C" int sun_link_monitor(sun_lattice* lat, void* stream, double out[6]) {
    if (!lat || !out) return fail(SUN_ERR_INVALID, "sun_link_monitor: NULL argument");
    if (!lat->hres) return fail(SUN_ERR_STATE, "sun_link_monitor: lattice has no host result buffer");
    hipStream_t s = pick_stream(lat, stream);
    SUN_TRY(link_monitor_to(lat, s, lat->dres));
    HIP_TRY(hipMemcpyAsync(lat->hres, lat->dres, 3 * sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) {
        out[2 * i] = lat->hres[i].x;
        out[2 * i + 1] = lat->hres[i].y;
    }
    return SUN_OK;
}

extern "C" int sun_link_monitor_async(sun_lattice* lat, void* stream, double* dev_out) {
    if (!lat || !dev_out) return fail(SUN_ERR_INVALID, "sun_link_monitor_async: NULL argument");
    return link_monitor_to(lat, pick_stream(lat, stream), (double2*)dev_out);
}

extern "C" int sun_link_trace(sun_lattice* lat, void* stream, double out[2]) {
    if (!out) return fail(SUN_ERR_INVALID, "sun_link_trace: NULL argument");
    double m[6];
    SUN_TRY(sun_link_monitor(lat, stream, m));
    out[0] = m[0];
    out[1] = m[1];
    return SUN_OK;
}

extern "C" int sun_link_det(sun_lattice* lat, void* stream, double out[2]) {
    if (!out) return fail(SUN_ERR_INVALID, "sun_link_det: NULL argument");
    double m[6];
    SUN_TRY(sun_link_monitor(lat, stream, m));
    out[0] = m[2];
    out[1] = m[3];
    return SUN_OK;
}

// ---------------------------------------------------------------------------------
// host exchange
// ---------------------------------------------------------------------------------
extern "C" int sun_links_to_host(sun_lattice* lat, double* host, int layout) {
    if (!lat || !host) return fail(SUN_ERR_INVALID, "sun_links_to_host: NULL argument");
    if (layout != SUN_LAYOUT_CANONICAL) return fail(SUN_ERR_UNSUPPORTED, "layout %d", layout);
    const size_t n = 4 * (size_t)lat->g.V * lat->d.nc * lat->d.nc * 2;
    double* tmp = nullptr;
    hipStream_t s = lat->ctx->stream;
    HIP_TRY(hipMalloc(&tmp, n * sizeof(double)));
    int rc = SUN_OK;
    hipError_t e = lat->ops->to_canonical(lat->f, 4 * lat->g.V, tmp, s);
    if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(SUN_ERR_HIP, "sun_links_to_host: %s", hipGetErrorString(e));
    hipFree(tmp);
    return rc;
}

extern "C" int sun_links_from_host(sun_lattice* lat, const double* host, int layout) {
    if (!lat || !host) return fail(SUN_ERR_INVALID, "sun_links_from_host: NULL argument");
    if (layout != SUN_LAYOUT_CANONICAL) return fail(SUN_ERR_UNSUPPORTED, "layout %d", layout);
    const size_t n = 4 * (size_t)lat->g.V * lat->d.nc * lat->d.nc * 2;
    double* tmp = nullptr;
    hipStream_t s = lat->ctx->stream;
    HIP_TRY(hipMalloc(&tmp, n * sizeof(double)));
    int rc = SUN_OK;
    hipError_t e = hipMemcpyAsync(tmp, host, n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = lat->ops->from_canonical(lat->f, 4 * lat->g.V, tmp, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(SUN_ERR_HIP, "sun_links_from_host: %s", hipGetErrorString(e));
    hipFree(tmp);
    return rc;
}

// ---------------------------------------------------------------------------------
// streams / events
// ---------------------------------------------------------------------------------
extern "C" int sun_stream_create(void** stream) {
    if (!stream) return fail(SUN_ERR_INVALID, "sun_stream_create: NULL");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return SUN_OK;
}
extern "C" int sun_stream_destroy(void* stream) {
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return SUN_OK;
}
extern "C" int sun_stream_synchronize(void* stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return SUN_OK;
}
extern "C" int sun_event_create(void** ev) {
    if (!ev) return fail(SUN_ERR_INVALID, "sun_event_create: NULL");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *ev = (void*)e;
    return SUN_OK;
}
extern "C" int sun_event_destroy(void* ev) {
    HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return SUN_OK;
}
extern "C" int sun_event_record(void* ev, void* stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return SUN_OK;
}
extern "C" int sun_event_elapsed_ms(void* a, void* b, float* ms) {
    if (!ms) return fail(SUN_ERR_INVALID, "sun_event_elapsed_ms: NULL");
    HIP_TRY(hipEventSynchronize((hipEvent_t)b));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return SUN_OK;
}
