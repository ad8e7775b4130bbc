/*
 * sun.hpp -- C++ facade over the C ABI (sun.h) with the class shapes of the
 * reference's generation API (namespace CULQCD), so a CULQCD program such as
 *
 *     initCULQCD(0, SUMMARIZE, TUNE_YES);
 *     SETPARAMS(false, 6.0, 32, 32, 32, 32, true);
 *     gauge U(SOA12, Device, PARAMS::Volume * 4, true);
 *     RNG rng; rng.Init(1234); U.Init(rng);
 *     HeatBath<float> hb(U, rng); OverRelaxation<float> ovr(U);
 *     Reunitarize<float> ru(U); PlaquetteCUB<float> plaq(U); OnePolyakovLoop<float> poly(U);
 *     for (...) { hb.Run(); for (k < 4) ovr.Run(); ru.Run(); plaq.Run(); poly.Run(); }
 *     U.Release(); EndCULQCD(0);
 *
 * compiles against this header after `using namespace sun::culqcd;`.
 *
 * Reference interfaces mirrored (file:line in the reference tree):
 *   initCULQCD / EndCULQCD         include/comm_mpi.h:43-45
 *   SETPARAMS, PARAMS::            include/constants.h:53-128
 *   _gauge<Real> (gauge)           include/gaugearray.h:88-465
 *   RNG                            include/random.h:42-101
 *   HeatBath<Real>                 include/monte/monte.h:17-75
 *   OverRelaxation<Real>           include/monte/ovr.h:20
 *   Reunitarize<Real>              include/reunitarize.h:22-60
 *   PlaquetteCUB<Real>             include/meas/plaquette.h:120-176
 *   OnePolyakovLoop<Real>          include/meas/polyakovloop.h:15-73
 *   SaveBin_Gauge / ReadBin_Gauge  include/io_gauge.h, src/io_gauge.cpp:429-700
 *   Sigma_g_plus, CalcWLOPs_A0,    include/meas/wloopex.h:13-36, src/wl/calcop_dg_A0.cu,
 *   CalcWilsonLoop_A0              src/wl/wilsonloop_dg_A0.cu
 *   CalcWLOPs_dg_33,               src/wl/calcop_dg_33.cu, src/wl/wilsonloop_dg.cu,
 *   CalcWilsonLoop_dg              include/meas/wlop.h
 *
 * Differences, on purpose: errors throw sun::culqcd::Error (a std::runtime_error
 * carrying sun_last_error()) instead of exiting; SETPARAMS only records the
 * geometry and beta used by gauge objects created afterwards (a process may hold
 * lattices of several geometries); N is a run-time property (NCOLORS below, default
 * 3) instead of a compile-time macro; the RNG is counter-based, so RNG holds only
 * the seed and HeatBath advances its own sweep counter.
 */
#ifndef SUN_HPP
#define SUN_HPP

#include <complex>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "sun.h"

namespace sun {
namespace culqcd {

class Error : public std::runtime_error {
   public:
    explicit Error(const std::string& what) : std::runtime_error(what) {}
};

inline void check(int rc, const char* what) {
    if (rc != SUN_OK) throw Error(std::string(what) + ": " + sun_last_error());
}

typedef enum { SILENT = SUN_SILENT, SUMMARIZE = SUN_SUMMARIZE, VERBOSE = SUN_VERBOSE,
               DEBUG_VERBOSE = SUN_DEBUG_VERBOSE } Verbosity;
typedef enum { TUNE_NO = SUN_TUNE_NO, TUNE_YES = SUN_TUNE_YES } TuneMode;
typedef enum { SOA = 0, SOA12 = 1, SOA12A = 2, SOA8 = 3 } ArrayType;  /* modes.h */
typedef enum { Host, Device } ReadMode;

/* process-wide context and current geometry (the reference's globals) */
struct Runtime {
    sun_ctx* ctx = nullptr;
    int nc = 3;
};
inline Runtime& runtime() {
    static Runtime r;
    return r;
}

namespace PARAMS {
inline double Beta = 0.0;
inline int Grid[4] = {0, 0, 0, 0};
inline int Volume = 0;
inline int HalfVolume = 0;
inline int size = 0;
inline bool UseTex = false;
}  // namespace PARAMS

/* initCULQCD(gpuid, verbosity, tune)   include/comm_mpi.h:43 */
inline void initCULQCD(int gpuid, Verbosity v = SILENT, TuneMode t = TUNE_NO) {
    if (runtime().ctx) return;
    check(sun_init(gpuid, (int)v, (int)t, &runtime().ctx), "initCULQCD");
}
/* EndCULQCD(status): releases the context; does not exit   include/comm_mpi.h:45 */
inline void EndCULQCD(int /*status*/) {
    if (runtime().ctx) {
        sun_finalize(runtime().ctx);
        runtime().ctx = nullptr;
    }
}
/* number of colours of lattices created from now on (compile-time NCOLORS in the reference) */
inline void SetNColors(int nc) { runtime().nc = nc; }

/* SETPARAMS(usetex, beta, nx, ny, nz, nt, verbose)   include/constants.h:128 */
inline void SETPARAMS(bool usetex, double beta, int nx, int ny, int nz, int nt, bool /*verbose*/ = false) {
    PARAMS::UseTex = usetex;
    PARAMS::Beta = beta;
    PARAMS::Grid[0] = nx;
    PARAMS::Grid[1] = ny;
    PARAMS::Grid[2] = nz;
    PARAMS::Grid[3] = nt;
    PARAMS::Volume = nx * ny * nz * nt;
    PARAMS::HalfVolume = PARAMS::Volume / 2;
    PARAMS::size = 4 * PARAMS::Volume;
}

/* RNG: holds the seed; draws are a function of (seed, sweep, parity, mu, site) */
class RNG {
   public:
    void Init(unsigned int seed) { seed_ = seed; }
    void Release() {}
    unsigned int& Seed() { return seed_; }

   private:
    unsigned int seed_ = 0;
};

/* _gauge<Real>: a shallow handle of one device link field, released explicitly */
template <class Real>
class _gauge {
   public:
    _gauge() {}
    _gauge(ArrayType type, ReadMode mode, int size = 0, bool evenodd = true, bool border = false) {
        (void)size;
        (void)border;
        if (mode != Device) throw Error("_gauge: only Device arrays are supported");
        // Device storage is always even/odd ordered; a lexicographic ("normal order")
        // array, which the reference's smearing and Wilson loops require
        // (ape.cu:202, wilsonloop.cu:504), differs only in its host exchange order.
        evenodd_ = evenodd;
        if (!runtime().ctx) throw Error("_gauge: call initCULQCD first");
        sun_lattice_desc d;
        d.nc = runtime().nc;
        for (int i = 0; i < 4; i++) d.dims[i] = PARAMS::Grid[i];
        d.beta = PARAMS::Beta;
        d.storage = type == SOA12 ? SUN_STORE_12 : type == SOA8 ? SUN_STORE_8 : SUN_STORE_18;
        if (type == SOA12A) throw Error("_gauge: SOA12A is not supported");
        d.precision = (int)sizeof(Real);
        check(sun_lattice_create(runtime().ctx, &d, &lat_), "_gauge::Allocate");
        type_ = type;
    }
    void Release() {
        if (lat_) sun_lattice_destroy(lat_);
        lat_ = nullptr;
    }
    /* Init(): cold start; Init(RNG): (Haar) hot start    gaugearray.h:765-791 */
    void Init() { check(sun_cold_start(lat_, nullptr), "_gauge::Init"); }
    void Init(RNG& r) { check(sun_hot_start(lat_, r.Seed(), nullptr), "_gauge::Init(RNG)"); }
    void Copy(_gauge<Real>& from) { check(sun_lattice_copy(lat_, from.lat_, nullptr), "_gauge::Copy"); }
    ArrayType Type() const { return type_; }
    bool EvenOdd() const { return evenodd_; }
    int Size() const { return PARAMS::size; }
    sun_lattice* handle() const { return lat_; }
    /* FP64 host copy double[4][V][N][N][2], sites in even/odd order (index.h:133-141)
       or, for a normal-order array, lexicographic order */
    void ToHost(double* host) const {
        if (evenodd_) {
            check(sun_links_to_host(lat_, host, SUN_LAYOUT_CANONICAL), "_gauge::ToHost");
            return;
        }
        std::vector<double> eo(host_reals());
        check(sun_links_to_host(lat_, eo.data(), SUN_LAYOUT_CANONICAL), "_gauge::ToHost");
        permute(eo.data(), host, true);
    }
    void FromHost(const double* host) {
        if (evenodd_) {
            check(sun_links_from_host(lat_, host, SUN_LAYOUT_CANONICAL), "_gauge::FromHost");
            return;
        }
        std::vector<double> eo(host_reals());
        permute(host, eo.data(), false);
        check(sun_links_from_host(lat_, eo.data(), SUN_LAYOUT_CANONICAL), "_gauge::FromHost");
    }

   private:
    size_t host_reals() const { return (size_t)4 * PARAMS::Volume * 2 * runtime().nc * runtime().nc; }
    // to_lex: even/odd src -> lexicographic dst; else the reverse
    void permute(const double* src, double* dst, bool to_lex) const {
        const size_t V = (size_t)PARAMS::Volume, m = (size_t)2 * runtime().nc * runtime().nc;
        const int* X = PARAMS::Grid;
        for (size_t lex = 0; lex < V; lex++) {
            const size_t x0 = lex % X[0], x1 = (lex / X[0]) % X[1], x2 = (lex / ((size_t)X[0] * X[1])) % X[2],
                         x3 = lex / ((size_t)X[0] * X[1] * X[2]);
            const size_t eo = lex / 2 + ((x0 + x1 + x2 + x3) & 1) * (V / 2);
            for (int mu = 0; mu < 4; mu++) {
                const size_t a = (mu * V + eo) * m, b = (mu * V + lex) * m;
                for (size_t k = 0; k < m; k++) {
                    if (to_lex)
                        dst[b + k] = src[a + k];
                    else
                        dst[a + k] = src[b + k];
                }
            }
        }
    }
    sun_lattice* lat_ = nullptr;
    ArrayType type_ = SOA;
    bool evenodd_ = true;
};
typedef _gauge<float> gauges;
typedef _gauge<double> gauged;

/* HeatBath<Real>(gauge, RNG).Run(): one sweep, then the sweep counter advances */
template <class Real>
class HeatBath {
   public:
    HeatBath(_gauge<Real>& U, RNG& r) : U_(U), rng_(r) {}
    void Run(void* stream = nullptr) { check(sun_heatbath(U_.handle(), rng_.Seed(), sweep_++, stream), "HeatBath"); }
    uint32_t Sweep() const { return sweep_; }
    void SetSweep(uint32_t s) { sweep_ = s; }

   private:
    _gauge<Real>& U_;
    RNG& rng_;
    uint32_t sweep_ = 0;
};

template <class Real>
class OverRelaxation {
   public:
    explicit OverRelaxation(_gauge<Real>& U) : U_(U) {}
    void Run(void* stream = nullptr) { check(sun_overrelax(U_.handle(), stream), "OverRelaxation"); }

   private:
    _gauge<Real>& U_;
};

template <class Real>
class Reunitarize {
   public:
    explicit Reunitarize(_gauge<Real>& U) : U_(U) {}
    void Run(void* stream = nullptr) { check(sun_reunitarize(U_.handle(), stream), "Reunitarize"); }

   private:
    _gauge<Real>& U_;
};

/* PlaquetteCUB<Real>(gauge).Run() -> complex(spatial, temporal) */
template <class Real>
class PlaquetteCUB {
   public:
    explicit PlaquetteCUB(_gauge<Real>& U) : U_(U) {}
    std::complex<double> Run(void* stream = nullptr) {
        double out[2];
        check(sun_plaquette(U_.handle(), stream, out), "PlaquetteCUB");
        value_ = std::complex<double>(out[0], out[1]);
        return value_;
    }
    std::complex<double> Value() const { return value_; }

   private:
    _gauge<Real>& U_;
    std::complex<double> value_;
};

template <class Real>
class OnePolyakovLoop {
   public:
    explicit OnePolyakovLoop(_gauge<Real>& U) : U_(U) {}
    std::complex<double> Run(void* stream = nullptr) {
        double out[2];
        check(sun_polyakov(U_.handle(), stream, out), "OnePolyakovLoop");
        value_ = std::complex<double>(out[0], out[1]);
        return value_;
    }
    std::complex<double> Value() const { return value_; }

   private:
    _gauge<Real>& U_;
    std::complex<double> value_;
};

/* GaugeTraceCUB<Real>(gauge).Run() -> mean Tr U / N (linktrsum.h:24-55);
   GaugeTrace<Real>(gauge, sum) takes the reference's device scratch array and ignores it */
template <class Real>
class GaugeTraceCUB {
   public:
    explicit GaugeTraceCUB(_gauge<Real>& U) : U_(U) {}
    std::complex<double> Run(void* stream = nullptr) {
        double out[2];
        check(sun_link_trace(U_.handle(), stream, out), "GaugeTrace");
        value_ = std::complex<double>(out[0], out[1]);
        return value_;
    }
    std::complex<double> Value() const { return value_; }
    void printValue() const { std::printf("Link Trace Sum: < %.12e : %.12e >\n", value_.real(), value_.imag()); }

   private:
    _gauge<Real>& U_;
    std::complex<double> value_;
};
template <class Real>
class GaugeTrace : public GaugeTraceCUB<Real> {
   public:
    GaugeTrace(_gauge<Real>& U, void* /*sum*/) : GaugeTraceCUB<Real>(U) {}
};

/* GaugeDetCUB<Real>(gauge).Run() -> mean det U (linkdetsum.h:23-66); GaugeDet alike */
template <class Real>
class GaugeDetCUB {
   public:
    explicit GaugeDetCUB(_gauge<Real>& U) : U_(U) {}
    std::complex<double> Run(void* stream = nullptr) {
        double out[2];
        check(sun_link_det(U_.handle(), stream, out), "GaugeDet");
        value_ = std::complex<double>(out[0], out[1]);
        return value_;
    }
    std::complex<double> Value() const { return value_; }
    void printValue() const {
        std::printf("Link Determinant Sum: < %.12e : %.12e >\n", value_.real(), value_.imag());
    }

   private:
    _gauge<Real>& U_;
    std::complex<double> value_;
};
template <class Real>
class GaugeDet : public GaugeDetCUB<Real> {
   public:
    explicit GaugeDet(_gauge<Real>& U) : GaugeDetCUB<Real>(U) {}
};

/* The measurement chain of the reference's test.cpp:170-230 (smearing, multihit,
   Wilson loops) on a gauge array of either ordering. */
/* ApplyAPEinSpace<Real>(gauge, w, steps, nhits, tol)   ape.cu:198-230 */
template <class Real>
void ApplyAPEinSpace(_gauge<Real>& U, Real w, int steps, int nhits, Real tol) {
    check(sun_ape_smear(U.handle(), (double)w, steps, nhits, (double)tol, 0, nullptr), "ApplyAPEinSpace");
}
/* ApplyAPEinTime<Real>(gauge, w, steps, nhits, tol)   ape.cu:236-264 */
template <class Real>
void ApplyAPEinTime(_gauge<Real>& U, Real w, int steps, int nhits, Real tol) {
    check(sun_ape_smear(U.handle(), (double)w, steps, nhits, (double)tol, 1, nullptr), "ApplyAPEinTime");
}
/* ApplyStoutinSpace<Real>(gauge, w, steps) / ApplyStoutinTime   stout.cu:687-766 (SU(3)) */
template <class Real>
void ApplyStoutinSpace(_gauge<Real>& U, Real w, int steps) {
    check(sun_stout_smear(U.handle(), (double)w, steps, 0, nullptr), "ApplyStoutinSpace");
}
template <class Real>
void ApplyStoutinTime(_gauge<Real>& U, Real w, int steps) {
    check(sun_stout_smear(U.handle(), (double)w, steps, 1, nullptr), "ApplyStoutinTime");
}
/* ParamHYP (smear.h): alpha1..3, defaults 0.75, 0.6, 0.3 */
struct ParamHYP {
    float alpha1 = 0.75f, alpha2 = 0.6f, alpha3 = 0.3f;
    ParamHYP() {}
    ParamHYP(float a1, float a2, float a3) : alpha1(a1), alpha2(a2), alpha3(a3) {}
    void set(float a1, float a2, float a3) {
        alpha1 = a1;
        alpha2 = a2;
        alpha3 = a3;
    }
};
/* ApplyHYPinSpace<Real>(gauge, steps, nhits, tol, ParamHYP) / ApplyHYPinTime   hyp.cu:277-368 */
template <class Real>
void ApplyHYPinSpace(_gauge<Real>& U, int steps, int nhits, Real tol, ParamHYP hyp) {
    check(sun_hyp_smear(U.handle(), hyp.alpha1, hyp.alpha2, hyp.alpha3, steps, nhits, (double)tol, 0, nullptr),
          "ApplyHYPinSpace");
}
template <class Real>
void ApplyHYPinTime(_gauge<Real>& U, int steps, int nhits, Real tol, ParamHYP hyp) {
    check(sun_hyp_smear(U.handle(), hyp.alpha1, hyp.alpha2, hyp.alpha3, steps, nhits, (double)tol, 1, nullptr),
          "ApplyHYPinTime");
}
/* ApplyMultiHit(in, out, RNG, nhit)   multihit.cu:291-333; out must be SOA (18-real) */
template <class Real>
void ApplyMultiHit(_gauge<Real>& in, _gauge<Real>& out, RNG& rng, int nhit) {
    check(sun_multihit(in.handle(), out.handle(), rng.Seed(), nhit, nullptr), "ApplyMultiHit");
}
/* ApplyMultiHitSpace(in, out, RNG, nhit)   multihitsp.cu:246-266 (the RNG is unused) */
template <class Real>
void ApplyMultiHitSpace(_gauge<Real>& in, _gauge<Real>& out, RNG& rng, int nhit) {
    (void)rng;
    check(sun_multihit_space(in.handle(), out.handle(), nhit, nullptr), "ApplyMultiHitSpace");
}
/* ApplyMultiHitExt(in, out, RNG, nhit)   multihitext.cu:449-456; out must be SOA (18-real) */
template <class Real>
void ApplyMultiHitExt(_gauge<Real>& in, _gauge<Real>& out, RNG& rng, int nhit) {
    check(sun_multihit_ext(in.handle(), out.handle(), rng.Seed(), nhit, nullptr), "ApplyMultiHitExt");
}
/* WilsonLoop<Real>(gauge, res, Rmax, Tmax)   wilsonloop.cu:494-516:
   res[it + r * (Tmax + 1)] = W(r, it), r <= Rmax, it <= Tmax */
template <class Real>
void WilsonLoop(_gauge<Real>& U, std::complex<double>* res, int radius, int Tmax) {
    std::vector<double> out((size_t)2 * (radius + 1) * (Tmax + 1));
    check(sun_wilson_loops(U.handle(), radius, Tmax, out.data(), nullptr), "WilsonLoop");
    for (size_t i = 0; i < out.size() / 2; i++) res[i] = std::complex<double>(out[2 * i], out[2 * i + 1]);
}
/* CWilsonLoop: the reference's lower-memory variant (wilsonloop.cu:420-468), same values */
template <class Real>
void CWilsonLoop(_gauge<Real>& U, std::complex<double>* res, int radius, int Tmax) {
    WilsonLoop<Real>(U, res, radius, Tmax);
}

/* Sigma_g_plus<Real>(Rmax, Tmax)   wloopex.h:13-36: the A0 operator basis (opN = 10
   operators, totalOpN = 100 pairs) with its correlator accumulator.  The device objects
   are created on the first CalcWLOPs_A0 with the gauge field it is given.  The
   reference's test clears and fetches the accumulator with raw cudaMemset / cudaMemcpy
   on arg.wloop (test.cpp:280, :285); here Clear() and Fetch() (which fills wloop_h,
   totalOpN (Tmax + 1) values, pair i opN + j at [it + pair (Tmax + 1)]). */
template <class Real>
class Sigma_g_plus {
   public:
    int Rmax, Tmax, opN, totalOpN;
    std::vector<Real> wloop_h;
    size_t wloop_size;
    Sigma_g_plus(int Rmax_, int Tmax_) : Rmax(Rmax_), Tmax(Tmax_), opN(10), totalOpN(100) {
        wloop_h.assign((size_t)totalOpN * (Tmax + 1), Real(0));
        wloop_size = wloop_h.size() * sizeof(Real);
    }
    ~Sigma_g_plus() {
        if (h_) sun_wlop_destroy(h_);
    }
    Sigma_g_plus(const Sigma_g_plus&) = delete;
    Sigma_g_plus& operator=(const Sigma_g_plus&) = delete;
    void Clear() {
        if (h_) check(sun_wlop_clear(h_, nullptr), "Sigma_g_plus::Clear");
    }
    void Fetch() {
        std::vector<double> v(wloop_h.size(), 0.0);
        if (h_) check(sun_wlop_read(h_, v.data(), nullptr), "Sigma_g_plus::Fetch");
        for (size_t i = 0; i < v.size(); i++) wloop_h[i] = (Real)v[i];
    }
    sun_wlop* attach(sun_lattice* lat) {
        if (!h_) check(sun_wlop_create(lat, SUN_WLOP_A0, Tmax, &h_), "Sigma_g_plus");
        return h_;
    }

   private:
    sun_wlop* h_ = nullptr;
};
/* CalcWLOPs_A0(gauge, &arg, radius, mu)   calcop_dg_A0.cu:307-343 */
template <class Real>
void CalcWLOPs_A0(_gauge<Real>& U, Sigma_g_plus<Real>* arg, int radius, int mu) {
    check(sun_wlop_fields(arg->attach(U.handle()), radius, mu, nullptr), "CalcWLOPs_A0");
}
/* CalcWilsonLoop_A0(gauge, &arg, radius, mu)   wilsonloop_dg_A0.cu:219-254 */
template <class Real>
void CalcWilsonLoop_A0(_gauge<Real>& U, Sigma_g_plus<Real>* arg, int radius, int mu) {
    check(sun_wlop_correlate(arg->attach(U.handle()), radius, mu, nullptr), "CalcWilsonLoop_A0");
}
/* The fieldOp of CalcWLOPs_dg_33 (calcop_dg_33.cu:966-987): the 25 operators of the
   33 family and the accumulator that CalcWilsonLoop_dg adds to. */
template <class Real>
class WLOpField33 {
   public:
    WLOpField33(_gauge<Real>& U, int Tmax) : Tmax(Tmax) {
        check(sun_wlop_create(U.handle(), SUN_WLOP_33, Tmax, &h_), "WLOpField33");
    }
    ~WLOpField33() { sun_wlop_destroy(h_); }
    WLOpField33(const WLOpField33&) = delete;
    WLOpField33& operator=(const WLOpField33&) = delete;
    sun_wlop* handle() const { return h_; }
    void Clear() { check(sun_wlop_clear(h_, nullptr), "WLOpField33::Clear"); }
    int Tmax;

   private:
    sun_wlop* h_ = nullptr;
};
template <class Real>
void CalcWLOPs_dg_33(_gauge<Real>& U, WLOpField33<Real>& fieldOp, int radius, int mu) {
    (void)U;
    check(sun_wlop_fields(fieldOp.handle(), radius, mu, nullptr), "CalcWLOPs_dg_33");
}
/* CalcWilsonLoop_dg(gauge, fieldOp, wloop, radius, Tmax, mu, opN)   wilsonloop_dg.cu:227-266:
   the reference reads its component table from __constant__ DEV_Ops / DEV_OpComps /
   DEV_OpPos, which no reference file sets; here the table is an argument.  Adds this
   (radius, mu) to the accumulator and returns the accumulated component-block pairs in
   wloop[it + idOp (Tmax + 1)]. */
template <class Real>
void CalcWilsonLoop_dg(_gauge<Real>& U, WLOpField33<Real>& fieldOp, Real* wloop, int radius, int mu,
                       const std::vector<int>& comps, const std::vector<int>& pos) {
    (void)U;
    check(sun_wlop_correlate(fieldOp.handle(), radius, mu, nullptr), "CalcWilsonLoop_dg");
    size_t n = 0;
    for (int c : comps) n += (size_t)c * c;
    std::vector<double> v(n * (fieldOp.Tmax + 1));
    check(sun_wlop_read_blocks(fieldOp.handle(), (int)comps.size(), comps.data(), pos.data(), v.data(), nullptr),
          "CalcWilsonLoop_dg");
    for (size_t i = 0; i < v.size(); i++) wloop[i] = (Real)v[i];
}

/* SaveBin_Gauge<Real, RealSaveConf>(gauge, filename, withheader) / ReadBin_Gauge:
 * the reference's binary configuration layout (io_gauge.cpp:429-583, :605-700);
 * RealSaveConf is the precision of the reals in the file. */
template <class RealSaveConf, class Real>
void SaveBin_Gauge(_gauge<Real>& U, const std::string& filename, bool withheader = true) {
    check(sun_lattice_save(U.handle(), filename.c_str(), withheader ? SUN_FILE_REFERENCE : SUN_FILE_REFERENCE_NOHEADER,
                           (int)sizeof(RealSaveConf), 0, 0),
          "SaveBin_Gauge");
}
template <class RealSaveConf, class Real>
void ReadBin_Gauge(_gauge<Real>& U, const std::string& filename, bool withheader = true) {
    check(sun_lattice_load(U.handle(), filename.c_str(), withheader ? SUN_FILE_REFERENCE : SUN_FILE_REFERENCE_NOHEADER,
                           (int)sizeof(RealSaveConf), nullptr, nullptr),
          "ReadBin_Gauge");
}

}  // namespace culqcd
}  // namespace sun

#endif /* SUN_HPP */
