// A CULQCD-style generation program written against include/sun.hpp (the C++
// facade of the reference's class API): SU(3), beta = 6.0, Haar hot start, FP64
// 18-real links, N iterations of 1 heatbath + 4 overrelaxation sweeps with the
// plaquette and Polyakov loop after each.  Prints one line per iteration
// ("iter plaq_s plaq_t poly_re poly_im", %.17g) and writes the final links
// (canonical double[4][V][3][3][2]) to the file given as the last argument.
// Usage: facade_demo X Y Z T seed iterations out.bin
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sun.hpp"

using namespace sun::culqcd;

int main(int argc, char** argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s X Y Z T seed iterations out.bin\n", argv[0]);
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
        U.Release();
        EndCULQCD(0);
    } catch (const Error& e) {
        fprintf(stderr, "facade_demo: %s\n", e.what());
        EndCULQCD(1);
        return 1;
    }
    return 0;
}
