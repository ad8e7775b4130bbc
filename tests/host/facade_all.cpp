// Instantiates every class and function template of the C++ facade (include/sun.hpp)
// for both precisions, so that each compiles against the C ABI; compiled (not run) by
// tests/test_abi.py.
#include "sun.hpp"

using namespace sun::culqcd;

template <class Real>
void use_all(_gauge<Real>& U, _gauge<Real>& out, RNG& rng) {
    HeatBath<Real> hb(U, rng);
    OverRelaxation<Real> ovr(U);
    Reunitarize<Real> ru(U);
    PlaquetteCUB<Real> plaq(U);
    OnePolyakovLoop<Real> poly(U);
    GaugeTraceCUB<Real> tr(U);
    GaugeTrace<Real> tr2(U, nullptr);
    GaugeDetCUB<Real> det(U);
    GaugeDet<Real> det2(U);
    hb.Run();
    ovr.Run();
    ru.Run();
    (void)plaq.Run();
    (void)poly.Run();
    (void)tr.Run();
    (void)tr2.Run();
    (void)det.Run();
    (void)det2.Run();
    tr.printValue();
    det.printValue();
    ApplyAPEinSpace<Real>(U, Real(0.2), 1, 10, Real(1e-10));
    ApplyAPEinTime<Real>(U, Real(0.2), 1, 10, Real(1e-10));
    ApplyStoutinSpace<Real>(U, Real(0.1), 1);
    ApplyStoutinTime<Real>(U, Real(0.1), 1);
    ParamHYP hyp;
    hyp.set(0.75f, 0.6f, 0.3f);
    ApplyHYPinSpace<Real>(U, 1, 10, Real(1e-10), hyp);
    ApplyHYPinTime<Real>(U, 1, 10, Real(1e-10), ParamHYP(1.f, 0.5f, 0.5f));
    ApplyMultiHit<Real>(U, out, rng, 4);
    ApplyMultiHitSpace<Real>(U, out, rng, 4);
    ApplyMultiHitExt<Real>(U, out, rng, 2);
    std::complex<double> res[4];
    WilsonLoop<Real>(out, res, 1, 1);
    CWilsonLoop<Real>(out, res, 1, 1);
    Sigma_g_plus<Real> arg(2, 1);
    arg.Clear();
    CalcWLOPs_A0<Real>(out, &arg, 2, 0);
    CalcWilsonLoop_A0<Real>(out, &arg, 2, 0);
    arg.Fetch();
    WLOpField33<Real> ops(out, 1);
    ops.Clear();
    CalcWLOPs_dg_33<Real>(out, ops, 2, 1);
    std::vector<Real> wl(2 * 2 * 2);
    CalcWilsonLoop_dg<Real>(out, ops, wl.data(), 2, 1, {1, 1}, {0, 4});
    SaveBin_Gauge<float>(U, "x.bin", true);
    ReadBin_Gauge<double>(U, "x.bin", false);
}

template void use_all<float>(gauges&, gauges&, RNG&);
template void use_all<double>(gauged&, gauged&, RNG&);
