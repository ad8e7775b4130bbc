// sun_meas.hip -- the reference's measurement chain on a resident lattice (host
// code): APE smearing (smear/ape.cu:198-264), stout smearing (smear/stout.cu:687-766),
// HYP smearing (smear/hyp.cu:277-368), spatial multihit (smear/multihitsp.cu:246-266),
// multihit of temporal links (smear/multihit.cu:291-333), extended multihit
// (smear/multihitext.cu:395-459) and Wilson loops (meas/wilsonloop.cu:494-516).
// Scratch (the ping-pong field of APE, the Wilson lines and partial sums) is
// allocated per call: these are measurements between generation steps, not part
// of the generation loop.
#include <hip/hip_runtime.h>

#include <vector>

#include "sun_internal.hpp"

using namespace sun;

namespace {

struct DeviceScratch {
    void* p = nullptr;
    ~DeviceScratch() {
        if (p) hipFree(p);
    }
};

int check_lattice(const sun_lattice* lat, const char* what) {
    if (!lat) return fail(SUN_ERR_INVALID, "%s: NULL lattice", what);
    if (lat->external) return fail(SUN_ERR_STATE, "%s: not available on a server job's lattice", what);
    return SUN_OK;
}

}  // namespace

extern "C" int sun_ape_smear(sun_lattice* lat, double w, int steps, int nhit, double tol, int time_dir,
                             void* stream) {
    SUN_TRY(check_lattice(lat, "sun_ape_smear"));
    if (steps < 0 || nhit < 0 || !(w >= 0.0) || (time_dir != 0 && time_dir != 1))
        return fail(SUN_ERR_INVALID, "sun_ape_smear: steps %d, nhit %d, w %g, time_dir %d", steps, nhit, w,
                    time_dir);
    if (steps == 0) return SUN_OK;
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    DeviceScratch tmp;
    HIP_TRY(hipMalloc(&tmp.p, lat->bytes));
    // ping-pong between the lattice and a copy: every direction of a step reads the
    // field of the previous step (ape.cu:209-214); the directions a step does not
    // smear are equal in both buffers
    HIP_TRY(hipMemcpyAsync(tmp.p, lat->buf, lat->bytes, hipMemcpyDeviceToDevice, s));
    const Field other = make_field(tmp.p, 4 * (size_t)lat->g.V, link_bytes(&lat->d));
    for (int st = 0; st < steps; st++) {
        const Field& src = (st & 1) ? other : lat->f;
        const Field& dst = (st & 1) ? lat->f : other;
        for (int mu = time_dir ? 3 : 0; mu < (time_dir ? 4 : 3); mu++)
            HIP_TRY(lat->ops->ape(src, dst, lat->g, mu, w, nhit, tol, s));
    }
    if (steps & 1) HIP_TRY(hipMemcpyAsync(lat->buf, tmp.p, lat->bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

extern "C" int sun_stout_smear(sun_lattice* lat, double rho, int steps, int time_dir, void* stream) {
    SUN_TRY(check_lattice(lat, "sun_stout_smear"));
    if (!lat->ops->stout) return fail(SUN_ERR_UNSUPPORTED, "sun_stout_smear: SU(3) only (N = %d)", lat->d.nc);
    if (steps < 0 || !(rho >= 0.0) || (time_dir != 0 && time_dir != 1))
        return fail(SUN_ERR_INVALID, "sun_stout_smear: steps %d, rho %g, time_dir %d", steps, rho, time_dir);
    if (steps == 0) return SUN_OK;
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    DeviceScratch tmp;
    HIP_TRY(hipMalloc(&tmp.p, lat->bytes));
    // the ping-pong of sun_ape_smear: every direction of a step reads the previous
    // step's field (stout.cu:697-705)
    HIP_TRY(hipMemcpyAsync(tmp.p, lat->buf, lat->bytes, hipMemcpyDeviceToDevice, s));
    const Field other = make_field(tmp.p, 4 * (size_t)lat->g.V, link_bytes(&lat->d));
    for (int st = 0; st < steps; st++) {
        const Field& src = (st & 1) ? other : lat->f;
        const Field& dst = (st & 1) ? lat->f : other;
        for (int mu = time_dir ? 3 : 0; mu < (time_dir ? 4 : 3); mu++)
            HIP_TRY(lat->ops->stout(src, dst, lat->g, mu, rho, s));
    }
    if (steps & 1) HIP_TRY(hipMemcpyAsync(lat->buf, tmp.p, lat->bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

extern "C" int sun_hyp_smear(sun_lattice* lat, double alpha1, double alpha2, double alpha3, int steps, int nhit,
                             double tol, int time_dir, void* stream) {
    SUN_TRY(check_lattice(lat, "sun_hyp_smear"));
    if (steps < 0 || nhit < 0 || (time_dir != 0 && time_dir != 1) || !(alpha1 >= 0.0) || !(alpha2 >= 0.0) ||
        !(alpha3 >= 0.0))
        return fail(SUN_ERR_INVALID, "sun_hyp_smear: steps %d, nhit %d, alphas %g %g %g, time_dir %d", steps, nhit,
                    alpha1, alpha2, alpha3, time_dir);
    if (steps == 0) return SUN_OK;
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    const size_t full = (size_t)2 * lat->d.nc * lat->d.nc * (size_t)lat->d.precision;  // 18-real record
    const size_t one = (size_t)lat->g.V * full;  // one pair's field
    DeviceScratch tmp, bar, til, tab;
    HIP_TRY(hipMalloc(&tmp.p, lat->bytes));
    HIP_TRY(hipMalloc(&bar.p, 12 * one));
    HIP_TRY(hipMalloc(&til.p, 12 * one));
    // 12 separate V-link fields per level (site-indexed plane views), as a device table
    Field views[24];
    for (int p = 0; p < 12; p++) {
        views[p] = make_field((char*)bar.p + p * one, (size_t)lat->g.V, full);
        views[12 + p] = make_field((char*)til.p + p * one, (size_t)lat->g.V, full);
    }
    HIP_TRY(hipMalloc(&tab.p, sizeof(views)));
    HIP_TRY(hipMemcpy(tab.p, views, sizeof(views), hipMemcpyHostToDevice));
    const Field* fb = (const Field*)tab.p;
    const Field* ft = fb + 12;
    // the ping-pong of sun_ape_smear (hyp.cu:291-313): each step reads the previous field
    HIP_TRY(hipMemcpyAsync(tmp.p, lat->buf, lat->bytes, hipMemcpyDeviceToDevice, s));
    const Field other = make_field(tmp.p, 4 * (size_t)lat->g.V, link_bytes(&lat->d));
    for (int st = 0; st < steps; st++) {
        const Field& src = (st & 1) ? other : lat->f;
        const Field& dst = (st & 1) ? lat->f : other;
        HIP_TRY(lat->ops->hyp(src, fb, ft, dst, lat->g, time_dir ? 3 : 0, time_dir ? 4 : 3, alpha1, alpha2, alpha3,
                              nhit, tol, s));
    }
    if (steps & 1) HIP_TRY(hipMemcpyAsync(lat->buf, tmp.p, lat->bytes, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

// the output of a multihit: another lattice, same context, N, precision and extents,
// 18-real storage (a mean of SU(N) matrices is not in SU(N))
static int check_average_output(const sun_lattice* in, const sun_lattice* out, const char* what) {
    SUN_TRY(check_lattice(in, what));
    SUN_TRY(check_lattice(out, what));
    if (in == out) return fail(SUN_ERR_INVALID, "%s: the output must be another lattice", what);
    if (out->d.storage != 18)
        return fail(SUN_ERR_INVALID, "%s: the output lattice must use 18-real storage (the averaged links are not "
                                     "in SU(N))", what);
    if (in->ctx != out->ctx || in->d.nc != out->d.nc || in->d.precision != out->d.precision)
        return fail(SUN_ERR_INVALID, "%s: input and output differ in context, N or precision", what);
    for (int i = 0; i < 4; i++)
        if (in->d.dims[i] != out->d.dims[i]) return fail(SUN_ERR_INVALID, "%s: the extents differ", what);
    return SUN_OK;
}

extern "C" int sun_multihit_space(sun_lattice* in, sun_lattice* out, int nhit, void* stream) {
    SUN_TRY(check_average_output(in, out, "sun_multihit_space"));
    if (nhit < 1) return fail(SUN_ERR_INVALID, "sun_multihit_space: nhit %d (the mean needs at least one)", nhit);
    hipStream_t s = pick_stream(in, stream);
    HIP_TRY(hipSetDevice(in->ctx->device));
    HIP_TRY(in->ops->multihit_space(in->f, out->f, in->g, nhit, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

extern "C" int sun_multihit(sun_lattice* in, sun_lattice* out, uint64_t seed, int nhit, void* stream) {
    SUN_TRY(check_average_output(in, out, "sun_multihit"));
    if (nhit < 0) return fail(SUN_ERR_INVALID, "sun_multihit: nhit %d", nhit);
    hipStream_t s = pick_stream(in, stream);
    HIP_TRY(hipSetDevice(in->ctx->device));
    HIP_TRY(in->ops->multihit(in->f, out->f, in->g, seed, nhit, in->d.beta / (double)in->d.nc, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

extern "C" int sun_multihit_ext(sun_lattice* in, sun_lattice* out, uint64_t seed, int nhit, void* stream) {
    SUN_TRY(check_average_output(in, out, "sun_multihit_ext"));
    if (nhit < 1) return fail(SUN_ERR_INVALID, "sun_multihit_ext: nhit %d (the mean needs at least one)", nhit);
    hipStream_t s = pick_stream(in, stream);
    HIP_TRY(hipSetDevice(in->ctx->device));
    HIP_TRY(in->ops->multihit_ext(in->f, out->f, in->g, seed, nhit, in->d.beta / (double)in->d.nc, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SUN_OK;
}

static int sun_wilson_loops_body(sun_lattice* lat, int rmax, int tmax, double* out, void* stream) {
    SUN_TRY(check_lattice(lat, "sun_wilson_loops"));
    if (!out) return fail(SUN_ERR_INVALID, "sun_wilson_loops: NULL output");
    const Geom& g = lat->g;
    if (rmax < 0 || tmax < 0 || rmax > g.X[0] || rmax > g.X[1] || rmax > g.X[2] || tmax > g.X[3])
        return fail(SUN_ERR_INVALID, "sun_wilson_loops: rmax %d, tmax %d outside the lattice", rmax, tmax);
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    const size_t full = (size_t)2 * lat->d.nc * lat->d.nc * (size_t)lat->d.precision;  // 18-real record
    const size_t nline = 3 * (size_t)g.V;
    const int nb = monitor_blocks(g);
    const size_t npart = (size_t)3 * (tmax + 1) * nb;
    const size_t nres = (size_t)(rmax + 1) * (tmax + 1);
    DeviceScratch lines, part, res;
    HIP_TRY(hipMalloc(&lines.p, nline * full));
    HIP_TRY(hipMalloc(&part.p, npart * sizeof(double2)));
    HIP_TRY(hipMalloc(&res.p, nres * sizeof(double2)));
    const Field lf = make_field(lines.p, nline, full);
    const double scale = 1.0 / (3.0 * lat->d.nc * (double)g.V);  // wilsonloop.cu:512
    for (int r = 0; r <= rmax; r++) {
        HIP_TRY(lat->ops->wl_lines(lat->f, lf, g, r, s));
        HIP_TRY(lat->ops->wl_loops(lat->f, lf, g, r, tmax, (double2*)part.p, s));
        HIP_TRY(launch_wl_finalize((const double2*)part.p, nb, tmax, scale, (double2*)res.p + (size_t)r * (tmax + 1),
                                   s));
    }
    std::vector<double2> h(nres);
    HIP_TRY(hipMemcpyAsync(h.data(), res.p, nres * sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t i = 0; i < nres; i++) {
        out[2 * i] = h[i].x;
        out[2 * i + 1] = h[i].y;
    }
    return SUN_OK;
}

extern "C" int sun_wilson_loops(sun_lattice* lat, int rmax, int tmax, double* out, void* stream) {
    return guard_alloc("sun_wilson_loops", [&] { return sun_wilson_loops_body(lat, rmax, tmax, out, stream); });
}

// The same kernels with the line field and the partial sums in the device's scratch
// array (its address from sun_init) and the results written straight to the caller's
// device buffer: no allocation and no host synchronisation.  The line field is set to
// the identity by the r = 0 pass and every partial is written before it is read, so
// nothing left in the array by an earlier call is used.
extern "C" int sun_wilson_loops_noalloc(sun_lattice* lat, int rmax, int tmax, double* dev_out, void* stream) {
    SUN_TRY(check_lattice(lat, "sun_wilson_loops_noalloc"));
    if (!dev_out || (reinterpret_cast<uintptr_t>(dev_out) & 15u) != 0)
        return fail(SUN_ERR_INVALID, "sun_wilson_loops_noalloc: dev_out must be a 16-byte aligned device pointer");
    const Geom& g = lat->g;
    if (rmax < 0 || tmax < 0 || rmax > g.X[0] || rmax > g.X[1] || rmax > g.X[2] || tmax > g.X[3])
        return fail(SUN_ERR_INVALID, "sun_wilson_loops_noalloc: rmax %d, tmax %d outside the lattice", rmax, tmax);
    if (!lat->ctx->scratch) return fail(SUN_ERR_STATE, "sun_wilson_loops_noalloc: no scratch array on this device");
    const size_t full = (size_t)2 * lat->d.nc * lat->d.nc * (size_t)lat->d.precision;  // 18-real record
    const size_t nline = 3 * (size_t)g.V;
    const int nb = monitor_blocks(g);
    const size_t line_bytes = (nline * full + 255) & ~(size_t)255;
    const size_t part_bytes = (size_t)3 * (tmax + 1) * nb * sizeof(double2);
    if (line_bytes + part_bytes > kNoAllocScratchBytes)
        return fail(SUN_ERR_BUDGET,
                    "sun_wilson_loops_noalloc: needs %zu bytes of scratch (line field %zu + partials %zu), the "
                    "array holds %zu; use sun_wilson_loops",
                    line_bytes + part_bytes, line_bytes, part_bytes, kNoAllocScratchBytes);
    hipStream_t s = pick_stream(lat, stream);
    HIP_TRY(hipSetDevice(lat->ctx->device));
    unsigned char* base = static_cast<unsigned char*>(lat->ctx->scratch);
    const Field lf = make_field(base, nline, full);
    double2* part = reinterpret_cast<double2*>(base + line_bytes);
    double2* res = reinterpret_cast<double2*>(dev_out);
    const double scale = 1.0 / (3.0 * lat->d.nc * (double)g.V);  // wilsonloop.cu:512
    for (int r = 0; r <= rmax; r++) {
        HIP_TRY(lat->ops->wl_lines(lat->f, lf, g, r, s));
        HIP_TRY(lat->ops->wl_loops(lat->f, lf, g, r, tmax, part, s));
        HIP_TRY(launch_wl_finalize(part, nb, tmax, scale, res + (size_t)r * (tmax + 1), s));
    }
    return SUN_OK;
}
