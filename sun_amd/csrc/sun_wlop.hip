// sun_wlop.hip -- operator-basis Wilson loops on a resident lattice (host code): the
// Sigma_g_plus measurement object (include/meas/wloopex.h:13-36) and the fieldOp /
// wloop pair of include/meas/wlop.h.  The A0 family (src/wl/calcop_dg_A0.cu, 10
// operators) and the 33 family (src/wl/calcop_dg_33.cu, 25 operators) build one
// 18-real operator field per slot; the correlator (src/wl/wilsonloop_dg_A0.cu,
// wilsonloop_dg.cu) accumulates every slot pair into an FP64 device array.  The
// reference's block reductions end in floating-point atomics (run-to-run order
// varies); here every (pair, T) is a fixed-order sum of per-workgroup partials.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "sun_internal.hpp"

using namespace sun;

struct sun_wlop {
    sun_lattice* lat = nullptr;
    int family = 0, nop = 0, tmax = 0;
    size_t lb = 0;          // bytes of one 18-real entry
    size_t stride = 0;      // bytes of one slot (V entries)
    void* ops = nullptr;    // nop slots
    double* partials = nullptr;
    int nb = 0;
    double* acc = nullptr;  // nop^2 (tmax + 1), (i nop + j)(tmax + 1) + T
    int r = -1, mu = -1;    // radius and direction of the current operator field
};

namespace {

int wlop_nops(int family) { return family == SUN_WLOP_A0 ? 10 : (family == SUN_WLOP_33 ? 25 : 0); }

void release(sun_wlop* h) {
    if (!h) return;
    if (h->ops) hipFree(h->ops);
    if (h->partials) hipFree(h->partials);
    if (h->acc) hipFree(h->acc);
    delete h;
}

int check_radius(const sun_wlop* h, int r, int mu, const char* what) {
    const Geom& g = h->lat->g;
    if (mu < 0 || mu > 2) return fail(SUN_ERR_INVALID, "%s: mu %d is not a spatial direction", what, mu);
    if (r < 1 || r > g.X[mu]) return fail(SUN_ERR_INVALID, "%s: radius %d outside 1..%d", what, r, g.X[mu]);
    return SUN_OK;
}

}  // namespace

extern "C" int sun_wlop_create(sun_lattice* lat, int family, int tmax, sun_wlop** out) {
    if (!out) return fail(SUN_ERR_INVALID, "sun_wlop_create: NULL handle pointer");
    *out = nullptr;
    if (!lat) return fail(SUN_ERR_INVALID, "sun_wlop_create: NULL lattice");
    if (lat->external) return fail(SUN_ERR_STATE, "sun_wlop_create: not available on a server job's lattice");
    const int nop = wlop_nops(family);
    if (!nop) return fail(SUN_ERR_INVALID, "sun_wlop_create: family %d (use SUN_WLOP_A0 or SUN_WLOP_33)", family);
    const Geom& g = lat->g;
    if (tmax < 0 || tmax > g.X[3]) return fail(SUN_ERR_INVALID, "sun_wlop_create: tmax %d outside 0..%d", tmax, g.X[3]);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    sun_wlop* h = new (std::nothrow) sun_wlop();
    if (!h) return fail(SUN_ERR_NOMEM, "sun_wlop_create: host allocation");
    h->lat = lat;
    h->family = family;
    h->nop = nop;
    h->tmax = tmax;
    h->lb = (size_t)2 * lat->d.nc * lat->d.nc * (size_t)lat->d.precision;
    h->stride = (size_t)g.V * h->lb;
    h->nb = monitor_blocks(g);  // the correlator's 256-site blocks
    const size_t nout = (size_t)nop * nop * (tmax + 1);
    HIP_TRY_CLEANUP(hipMalloc(&h->ops, (size_t)nop * h->stride), release(h));
    HIP_TRY_CLEANUP(hipMalloc((void**)&h->partials, nout * h->nb * sizeof(double)), release(h));
    HIP_TRY_CLEANUP(hipMalloc((void**)&h->acc, nout * sizeof(double)), release(h));
    HIP_TRY_CLEANUP(hipMemset(h->acc, 0, nout * sizeof(double)), release(h));
    *out = h;
    return SUN_OK;
}

extern "C" int sun_wlop_destroy(sun_wlop* h) {
    if (!h) return SUN_OK;
    hipSetDevice(h->lat->ctx->device);
    hipDeviceSynchronize();
    release(h);
    return SUN_OK;
}

extern "C" int sun_wlop_nops(const sun_wlop* h) { return h ? h->nop : 0; }

extern "C" int sun_wlop_fields(sun_wlop* h, int r, int mu, void* stream) {
    if (!h) return fail(SUN_ERR_INVALID, "sun_wlop_fields: NULL handle");
    SUN_TRY(check_radius(h, r, mu, "sun_wlop_fields"));
    sun_lattice* lat = h->lat;
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    const Field ops0 = make_field(h->ops, (size_t)lat->g.V, h->lb);
    HIP_TRY(lat->ops->wlop_fields(lat->f, ops0, h->stride, lat->g, h->family, r, mu, 2.0 * std::sqrt(2.0), s));
    h->r = r;
    h->mu = mu;
    return SUN_OK;
}

extern "C" int sun_wlop_correlate(sun_wlop* h, int r, int mu, void* stream) {
    if (!h) return fail(SUN_ERR_INVALID, "sun_wlop_correlate: NULL handle");
    SUN_TRY(check_radius(h, r, mu, "sun_wlop_correlate"));
    if (h->r < 0) return fail(SUN_ERR_STATE, "sun_wlop_correlate: no operator field yet (call sun_wlop_fields)");
    sun_lattice* lat = h->lat;
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    const Field ops0 = make_field(h->ops, (size_t)lat->g.V, h->lb);
    HIP_TRY(lat->ops->wlop_corr(lat->f, ops0, h->stride, lat->g, h->nop, r, mu, h->tmax, h->partials, s));
    HIP_TRY(launch_wlop_finalize(h->partials, h->nb, h->nop * h->nop * (h->tmax + 1), h->acc, s));
    return SUN_OK;
}

extern "C" int sun_wlop_clear(sun_wlop* h, void* stream) {
    if (!h) return fail(SUN_ERR_INVALID, "sun_wlop_clear: NULL handle");
    hipStream_t s = pick_stream(h->lat, stream);
    HIP_TRY(hipSetDevice(h->lat->ctx->device));
    HIP_TRY(hipMemsetAsync(h->acc, 0, (size_t)h->nop * h->nop * (h->tmax + 1) * sizeof(double), s));
    return SUN_OK;
}

extern "C" int sun_wlop_read(sun_wlop* h, double* out, void* stream) {
    if (!h || !out) return fail(SUN_ERR_INVALID, "sun_wlop_read: NULL argument");
    hipStream_t s = pick_stream(h->lat, stream);
    HIP_TRY(hipSetDevice(h->lat->ctx->device));
    HIP_TRY(hipMemcpyAsync(out, h->acc, (size_t)h->nop * h->nop * (h->tmax + 1) * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

static int sun_wlop_read_blocks_body(sun_wlop* h, int ncomp, const int* comps, const int* pos, double* out,
                                     void* stream) {
    if (!h || !out || ncomp < 0 || (ncomp > 0 && (!comps || !pos)))
        return fail(SUN_ERR_INVALID, "sun_wlop_read_blocks: bad arguments");
    for (int c = 0; c < ncomp; c++)
        if (comps[c] < 0 || pos[c] < 0 || pos[c] + comps[c] > h->nop)
            return fail(SUN_ERR_INVALID, "sun_wlop_read_blocks: component %d (%d operators from slot %d) outside %d slots",
                        c, comps[c], pos[c], h->nop);
    const int T1 = h->tmax + 1;
    std::vector<double> all((size_t)h->nop * h->nop * T1);
    SUN_TRY(sun_wlop_read(h, all.data(), stream));
    size_t id = 0;
    for (int c = 0; c < ncomp; c++)
        for (int a = 0; a < comps[c]; a++)
            for (int b = 0; b < comps[c]; b++, id++) {
                const int i = pos[c] + a, j = pos[c] + b;
                memcpy(out + id * T1, all.data() + (size_t)(i * h->nop + j) * T1, T1 * sizeof(double));
            }
    return SUN_OK;
}

extern "C" int sun_wlop_read_blocks(sun_wlop* h, int ncomp, const int* comps, const int* pos, double* out,
                                    void* stream) {
    return guard_alloc("sun_wlop_read_blocks",
                       [&] { return sun_wlop_read_blocks_body(h, ncomp, comps, pos, out, stream); });
}

static int sun_wlop_fields_to_host_body(sun_wlop* h, double* out) {
    if (!h || !out) return fail(SUN_ERR_INVALID, "sun_wlop_fields_to_host: NULL argument");
    if (h->r < 0) return fail(SUN_ERR_STATE, "sun_wlop_fields_to_host: no operator field yet");
    sun_lattice* lat = h->lat;
    HIP_TRY(hipSetDevice(lat->ctx->device));
    HIP_TRY(hipStreamSynchronize(pick_stream(lat, nullptr)));
    HIP_TRY(hipDeviceSynchronize());
    const size_t V = (size_t)lat->g.V, lb = h->lb;
    const int nreal = 2 * lat->d.nc * lat->d.nc;
    const bool f32 = lat->d.precision == SUN_PREC_F32;
    const size_t nfull = lb / 16, tail = lb % 16;
    std::vector<unsigned char> raw(h->stride);
    for (int k = 0; k < h->nop; k++) {
        HIP_TRY(hipMemcpy(raw.data(), (char*)h->ops + (size_t)k * h->stride, h->stride, hipMemcpyDeviceToHost));
        for (size_t l = 0; l < V; l++) {
            unsigned char rec[16 * kMaxPlanes];
            for (size_t c = 0; c < nfull; c++) memcpy(rec + 16 * c, raw.data() + c * V * 16 + l * 16, 16);
            if (tail) memcpy(rec + 16 * nfull, raw.data() + nfull * V * 16 + l * 8, 8);
            double* dst = out + ((size_t)k * V + l) * nreal;
            for (int q = 0; q < nreal; q++) {
                if (f32) {
                    float v;
                    memcpy(&v, rec + 4 * q, 4);
                    dst[q] = v;
                } else {
                    memcpy(dst + q, rec + 8 * q, 8);
                }
            }
        }
    }
    return SUN_OK;
}

extern "C" int sun_wlop_fields_to_host(sun_wlop* h, double* out) {
    return guard_alloc("sun_wlop_fields_to_host", [&] { return sun_wlop_fields_to_host_body(h, out); });
}

// test.cpp:273-298 with the operator basis of `family`: for r = 2..rmax the
// correlators of every slot pair summed over mu < 3, divided by 3 N V
static int sun_wilson_loops_ops_body(sun_lattice* lat, int family, int rmax, int tmax, double* out, void* stream) {
    if (!lat || !out) return fail(SUN_ERR_INVALID, "sun_wilson_loops_ops: NULL argument");
    const Geom& g = lat->g;
    if (rmax < 2 || rmax > g.X[0] || rmax > g.X[1] || rmax > g.X[2])
        return fail(SUN_ERR_INVALID, "sun_wilson_loops_ops: rmax %d outside 2..(smallest spatial extent)", rmax);
    sun_wlop* h = nullptr;
    SUN_TRY(sun_wlop_create(lat, family, tmax, &h));
    const size_t n = (size_t)h->nop * h->nop * (tmax + 1);
    const double norm = (double)g.V * 3.0 * lat->d.nc;  // test.cpp:290
    int rc = SUN_OK;
    for (int r = 2; r <= rmax && rc == SUN_OK; r++) {
        rc = sun_wlop_clear(h, stream);
        for (int mu = 0; mu < 3 && rc == SUN_OK; mu++) {
            rc = sun_wlop_fields(h, r, mu, stream);
            if (rc == SUN_OK) rc = sun_wlop_correlate(h, r, mu, stream);
        }
        if (rc == SUN_OK) rc = sun_wlop_read(h, out + (size_t)(r - 2) * n, stream);
        if (rc == SUN_OK)
            for (size_t k = 0; k < n; k++) out[(size_t)(r - 2) * n + k] /= norm;
    }
    sun_wlop_destroy(h);
    return rc;
}

extern "C" int sun_wilson_loops_ops(sun_lattice* lat, int family, int rmax, int tmax, double* out, void* stream) {
    return guard_alloc("sun_wilson_loops_ops",
                       [&] { return sun_wilson_loops_ops_body(lat, family, rmax, tmax, out, stream); });
}
