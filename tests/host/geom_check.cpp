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
    // every brick shape that brick_block accepts maps (workgroup, thread) onto the
    // sites of one parity exactly once, and the site's coordinates lie in its brick
    for (int b1 = 0; b1 <= 4; b1++)
        for (int b2 = 0; b2 <= 4; b2++)
            for (int b3 = 0; b3 <= 4; b3++) {
                const int block = sun::brick_block(g, b1, b2, b3, 1024);
                if (!block) continue;
                if (g.half % block) { printf("BRICK FAIL size %d %d %d\n", b1, b2, b3); return 1; }
                const int grid = g.half / block;
                std::vector<int> seen(g.half, 0);
                for (int b = 0; b < grid; b++)
                    for (int t = 0; t < block; t++) {
                        const int id = sun::brick_site(g, b, t, b1, b2, b3);
                        if (id < 0 || id >= g.half || seen[id]++) {
                            printf("BRICK FAIL %d %d %d b=%d t=%d\n", b1, b2, b3, b, t);
                            return 1;
                        }
                        int x[4];
                        sun::site_coords(g, id, 0, x);
                        const int lo1 = sun::brick_site(g, b, 0, b1, b2, b3);
                        int y[4];
                        sun::site_coords(g, lo1, 0, y);
                        if (x[1] - y[1] < 0 || x[1] - y[1] >= (1 << b1) || x[2] - y[2] < 0 ||
                            x[2] - y[2] >= (1 << b2) || x[3] - y[3] < 0 || x[3] - y[3] >= (1 << b3)) {
                            printf("BRICK FAIL shape %d %d %d b=%d t=%d\n", b1, b2, b3, b, t);
                            return 1;
                        }
                    }
                fprintf(stderr, "brick %d %d %d ok (block %d)\n", b1, b2, b3, block);
            }
    // the reduction workgroups cover every site once (threads past V in the linear
    // order are the only ones left over)
    {
        const int nb = sun::plaquette_blocks(g);
        std::vector<int> seen(g.V, 0);
        int outside = 0;
        for (int lb = 0; lb < nb; lb++)
            for (int t = 0; t < sun::kReduceBlock; t++) {
                const int idd = sun::reduce_site(g, lb, t);
                if (idd >= g.V && g.rb1 < 0) { outside++; continue; }
                if (idd < 0 || idd >= g.V || seen[idd]++) { printf("REDUCE FAIL lb=%d t=%d\n", lb, t); return 1; }
            }
        for (int i = 0; i < g.V; i++)
            if (!seen[i]) { printf("REDUCE MISSING %d\n", i); return 1; }
        fprintf(stderr, "reduce brick %d %d %d ok\n", g.rb1, g.rb2, g.rb3);
    }
    // the batched update brick, where one exists, is a shape brick_block accepts with
    // exactly 256 * spt sites, V/2 / (256 spt) bricks tile the parity exactly (spt
    // workgroups of 256 threads each), and the brick mapping is a bijection
    {
        int b1, b2, b3, spt;
        sun::batch_brick(g, b1, b2, b3, spt);
        const int L = g.X[0] / 2;
        const bool pow2 = L > 0 && L <= 256 && (256 % L) == 0;
        if (b1 >= 0) {
            const int sites = 256 * spt;
            if ((spt != 1 && spt != 3) || sun::brick_block(g, b1, b2, b3, sites) != sites || g.half % sites) {
                printf("BATCH BRICK FAIL %d %d %d spt %d\n", b1, b2, b3, spt);
                return 1;
            }
            std::vector<int> seen(g.half, 0);
            for (int lb = 0; lb < g.half / 256; lb++)
                for (int t = 0; t < 256; t++) {
                    const int B = lb / spt;
                    const int id = sun::brick_site(g, B, (lb - B * spt) * 256 + t, b1, b2, b3);
                    if (id < 0 || id >= g.half || seen[id]++) {
                        printf("BATCH BRICK MAP FAIL lb=%d t=%d\n", lb, t);
                        return 1;
                    }
                }
            fprintf(stderr, "batch brick %d %d %d spt %d ok\n", b1, b2, b3, spt);
        } else if (spt != 1) {
            printf("BATCH BRICK SPT WITHOUT BRICK %d\n", spt);
            return 1;
        } else if (pow2 && (long long)g.half % 256 == 0) {
            // no brick although X0/2 divides 256: only when no split of the rows over
            // x1, x2, x3 divides the extents
            for (int c1 = 0; c1 <= 8; c1++)
                for (int c2 = 0; c2 <= 8; c2++)
                    for (int c3 = 0; c3 <= 8; c3++)
                        if (sun::brick_block(g, c1, c2, c3, 256) == 256 && c1 + c2 + c3 <= 6) {
                            printf("BATCH BRICK MISSED %d %d %d\n", c1, c2, c3);
                            return 1;
                        }
        }
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
