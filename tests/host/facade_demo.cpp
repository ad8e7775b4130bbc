// A CULQCD-style generation program written against include/sun.hpp (the C++
// facade of the reference's class API): SU(3), beta = 6.0, Haar hot start, FP64
// 18-real links, N iterations of 1 heatbath + 4 overrelaxation sweeps with the
// plaquette and Polyakov loop after each.  Prints one line per iteration
// ("iter plaq_s plaq_t poly_re poly_im", %.17g) and writes the final links
// (canonical double[4][V][3][3][2]) to out.bin.  Then the measurement chain of the
// reference's test.cpp:170-230 on normal-order copies: APE smearing in space
// (w 0.2, 3 steps, 10 hits), multihit (seed + 1, 5 hits) into a second array and
// Wilson loops up to r = 2, T = 3, printed as "wloop r T re im" lines; the multihit
// array's lexicographic host copy goes to lex.bin.
// Usage: facade_demo X Y Z T seed iterations out.bin lex.bin
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sun.hpp"

using namespace sun::culqcd;

int main(int argc, char** argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s X Y Z T seed iterations out.bin lex.bin\n", argv[0]);
        return 2;
    }
    const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]), nt = atoi(argv[4]);
    const unsigned seed = (unsigned)atoi(argv[5]);
    const int iters = atoi(argv[6]);
    try {
        initCULQCD(0, SILENT, TUNE_NO);
        SETPARAMS(false, 6.0, nx, ny, nz, nt, false);
        gauged U(SOA, Device, PARAMS::Volume * 4, true);
        RNG rng;
        rng.Init(seed);
        U.Init(rng);
        HeatBath<double> hb(U, rng);
        OverRelaxation<double> ovr(U);
        PlaquetteCUB<double> plaq(U);
        OnePolyakovLoop<double> poly(U);
        for (int it = 0; it < iters; it++) {
            hb.Run();
            for (int k = 0; k < 4; k++) ovr.Run();
            std::complex<double> p = plaq.Run();
            std::complex<double> l = poly.Run();
            printf("%d %.17g %.17g %.17g %.17g\n", it, p.real(), p.imag(), l.real(), l.imag());
        }
        GaugeTraceCUB<double> trace(U);
        GaugeDet<double> det(U);
        std::complex<double> t = trace.Run(), d = det.Run();
        printf("monitor %.17g %.17g %.17g %.17g\n", t.real(), t.imag(), d.real(), d.imag());
        std::vector<double> links((size_t)4 * PARAMS::Volume * 18);
        U.ToHost(links.data());
        FILE* f = fopen(argv[7], "wb");
        if (!f) throw Error("cannot open output file");
        fwrite(links.data(), sizeof(double), links.size(), f);
        fclose(f);

        gauged BB(SOA, Device, PARAMS::Volume * 4, false);  // normal order, as test.cpp:173
        BB.Copy(U);
        ApplyAPEinSpace<double>(BB, 0.2, 3, 10, 1.e-10);
        gauged DD(SOA, Device, PARAMS::Volume * 4, false);
        RNG hits;
        hits.Init(seed + 1);
        ApplyMultiHit<double>(BB, DD, hits, 5);
        const int rmax = 2, tmax = 3;
        std::vector<std::complex<double>> res((rmax + 1) * (tmax + 1));
        WilsonLoop<double>(DD, res.data(), rmax, tmax);
        for (int r = 0; r <= rmax; r++)
            for (int it = 0; it <= tmax; it++)
                printf("wloop %d %d %.17g %.17g\n", r, it, res[it + r * (tmax + 1)].real(),
                       res[it + r * (tmax + 1)].imag());
        DD.ToHost(links.data());
        f = fopen(argv[8], "wb");
        if (!f) throw Error("cannot open output file");
        fwrite(links.data(), sizeof(double), links.size(), f);
        fclose(f);
        DD.Release();
        BB.Release();
        U.Release();
        EndCULQCD(0);
    } catch (const Error& e) {
        fprintf(stderr, "facade_demo: %s\n", e.what());
        EndCULQCD(1);
        return 1;
    }
    return 0;
}
