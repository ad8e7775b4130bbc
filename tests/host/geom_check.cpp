// Host-side check of the engine's device geometry (sun_field.hpp): prints, for every
// half-lattice site of both parities, its coordinates and the even/odd index of the
// eight +-1 neighbours and of x+mu-nu, so the test can compare them with the oracle
// and bound every index by V.  Usage: geom_check X Y Z T
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sun_field.hpp"
#include "sun_ops.hpp"

int main(int argc, char** argv) {
    if (argc != 5) return 2;
    int dims[4];
    for (int i = 0; i < 4; i++) dims[i] = atoi(argv[i + 1]);
    sun::Geom g = sun::make_geom(dims);
    for (int parity = 0; parity < 2; parity++)
        for (int id = 0; id < g.half; id++) {
            int x[4];
            int lex = sun::site_coords(g, id, parity, x);
            printf("%d %d %d %d %d %d", parity, id, x[0], x[1], x[2], x[3]);
            for (int mu = 0; mu < 4; mu++)
                for (int d = -1; d <= 1; d += 2) {
                    int y[4] = {x[0], x[1], x[2], x[3]};
                    int l2 = sun::lex_shift(g, lex, y, mu, d);
                    printf(" %d", sun::eo_of_lex(g, l2, parity ^ 1));
                }
            for (int mu = 0; mu < 4; mu++)
                for (int nu = 0; nu < 4; nu++) {
                    if (mu == nu) continue;
                    int y[4] = {x[0], x[1], x[2], x[3]};
                    int l2 = sun::lex_shift(g, lex, y, nu, -1);
                    l2 = sun::lex_shift(g, l2, y, mu, +1);
                    printf(" %d", sun::eo_of_lex(g, l2, parity));
                }
            printf("\n");
        }
    // FastDiv against hardware division over a range of divisors and numerators
    for (uint32_t d = 1; d <= 200; d++) {
        sun::FastDiv f = sun::FastDiv::make(d);
        for (uint32_t n = 0; n < 70000; n++)
            if (f.div(n) != n / d) { printf("FASTDIV FAIL %u %u\n", n, d); return 1; }
        for (uint32_t n = 0x7fffffffu - 70000; n <= 0x7fffffffu - 1; n++)
            if (f.div(n) != n / d) { printf("FASTDIV FAIL %u %u\n", n, d); return 1; }
    }
    // XCD remap of workgroups is a bijection of [0, n) for every grid size
    for (int n = 1; n <= 5000; n++) {
        std::vector<int> seen(n, 0);
        for (int b = 0; b < n; b++) {
            int v = sun::xcd_remap(b, n);
            if (v < 0 || v >= n || seen[v]++) { printf("REMAP FAIL %d %d\n", n, b); return 1; }
        }
    }
    return 0;
}
