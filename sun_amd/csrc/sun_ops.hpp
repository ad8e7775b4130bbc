// sun_ops.hpp -- the per-variant launcher table shared by the kernel translation
// units and the C-ABI layer.  One table exists for every (N, storage, precision)
// combination the engine supports; the C-ABI picks the table of a lattice once at
// creation.
#pragma once

#include <hip/hip_runtime.h>

#include "sun_field.hpp"

namespace sun {

// Launch shape chosen by the autotuner.
struct LaunchCfg {
    int block = 256;
    // dynamic LDS reserved per workgroup and not used by the kernel: caps the
    // workgroups resident per CU (160 KB of LDS per CU), i.e. the lattice volume in
    // flight per XCD, which decides whether neighbour links are still in L2 when
    // the next staple direction needs them
    int lds = 0;
    // brick mapping (brick_site) when b1 >= 0: log2 of the rows per workgroup along
    // x1, x2, x3; b1 < 0: linear mapping with `block` threads
    int b1 = -1, b2 = 0, b3 = 0;
    // 1: the software-pipelined overrelaxation kernel (where the variant has one)
    int piped = 0;
};

// Brick mapping of workgroups to lattice sites.  The sites of one parity form rows
// of L = X0/2 consecutive ids (fixed x1, x2, x3).  Workgroup b takes a brick of
// 2^b1 x 2^b2 x 2^b3 rows along (x1, x2, x3) -- L * 2^(b1+b2+b3) threads, row
// index x1-fastest inside the brick, bricks numbered x1-fastest -- so the sites
// whose staples share links sit in one workgroup, on one CU, at the same time.
// Returns the site id within the parity of thread t of workgroup b.
SUN_HD int brick_site(const Geom& g, int b, int t, int b1, int b2, int b3) {
    const int L = g.X[0] >> 1;
    const int row = (int)g.dxh.div((uint32_t)t);
    const int s = t - row * L;
    const int i1 = row & ((1 << b1) - 1);
    const int i2 = (row >> b1) & ((1 << b2) - 1);
    const int i3 = row >> (b1 + b2);
    const int n1 = g.X[1] >> b1, n2 = g.X[2] >> b2;
    const int B23 = b / n1, B1 = b - B23 * n1;
    const int B3 = B23 / n2, B2 = B23 - B3 * n2;
    const int x1 = (B1 << b1) + i1, x2 = (B2 << b2) + i2, x3 = (B3 << b3) + i3;
    return ((x3 * g.X[2] + x2) * g.X[1] + x1) * L + s;
}

constexpr int kReduceBlock = 256;   // threads of every reduction block

// site (0 <= idd < V, parity-major) of thread t of reduction workgroup lb of a
// lattice: workgroups 2k and 2k+1 take brick k of parity 0 and of parity 1 (whose
// links are each other's neighbours, so they run on the same XCD and share its L2);
// linear when the geometry has no reduction brick.  The single-lattice and the
// batched plaquette kernels share it, so their per-workgroup partial sums agree.
SUN_HD int reduce_site(const Geom& g, int lb, int t) {
    if (g.rb1 < 0) return lb * kReduceBlock + t;
    const int parity = lb & 1;
    return parity * g.half + brick_site(g, lb >> 1, t, g.rb1, g.rb2, g.rb3);
}

// threads per workgroup of a brick shape on this geometry, or 0 when the brick does
// not tile the lattice or the workgroup is not whole wavefronts within max_block
inline int brick_block(const Geom& g, int b1, int b2, int b3, int max_block) {
    if (b1 < 0 || b2 < 0 || b3 < 0 || b1 > 6 || b2 > 6 || b3 > 6) return 0;
    if ((g.X[1] % (1 << b1)) || (g.X[2] % (1 << b2)) || (g.X[3] % (1 << b3))) return 0;
    const long long block = (long long)(g.X[0] / 2) << (b1 + b2 + b3);
    if (block < 64 || block > max_block || (block % 64)) return 0;
    return (int)block;
}

// One lattice of a batched launch.  `first_block` is the prefix sum of the blocks
// of the preceding lattices; a block finds its lattice by binary search.  Each
// kind of batched launch (update, reunitarisation, plaquette, Polyakov loop) has
// its own table because the blocks per lattice differ.
struct BatchEntry {
    Field f;
    Geom g;
    uint64_t seed;
    double bnc;          // beta / N
    uint32_t sweep;      // heatbath sweep counter of this lattice
    int first_block;
    int nblocks;
    double2* partials;   // this lattice's per-block reduction scratch (nblocks entries)
    double scale;        // normalisation of the reduced sum
    double2* hist;       // observable history: {plaquette, polyakov} per iteration
    uint32_t hist_iter;  // next history row
    int ub1, ub2, ub3;   // update brick (ub1 < 0: none)
    int spt;             // 256-thread workgroups per update brick (the brick holds 256 spt sites)
};

// The update brick of a batched lattice: 256 * spt sites = (X0/2) * 2^k rows, taken by
// spt consecutive 256-thread workgroups (same XCD, side by side; every workgroup of a
// batch does the same work, so the XCDs' contiguous ranges stay balanced).  spt = 1
// where rows of X0/2 tile 256 sites: the k doublings go to x1, x3, x1, x3 (the tuned
// single-lattice shape at 32^4 is such a (2, 0, 2) brick), then x2, x2, then
// round-robin, each only where the extent stays divisible.  Else spt = 3 where rows
// tile 768 sites (X0/2 = 12: 24^4, 64 rows), doublings x1, x3, x1, x3, x1, x3 (8 x 1 x
// 8 rows at 24^4: in the cache-sized chunks of a batch the best of five shapes
// measured, 4 x 2 x 8 and 4 x 4 x 4 rows 1-2% slower).  b1 < 0 (spt = 1) when no brick
// tiles the lattice.
inline void batch_brick(const Geom& g, int& b1, int& b2, int& b3, int& spt) {
    b1 = -1;
    b2 = b3 = 0;
    spt = 1;
    const int L = g.X[0] / 2;
    if (L <= 0 || L > 768) return;
    int m = 0;
    if (256 % L == 0) m = 1;
    else if (768 % L == 0 && ((768 / L) & (768 / L - 1)) == 0) m = 3;
    if (!m) return;
    int k = 0;
    while ((L << k) < 256 * m) k++;
    static const int seq1[] = {1, 3, 1, 3, 2, 2, 1, 3, 2, 1, 3, 2, 1, 3, 2, 1, 3, 2};
    static const int seq3[] = {1, 3, 1, 3, 1, 3, 2, 2, 1, 3, 2, 1, 3, 2, 1, 3, 2, 2};
    const int* seq = m == 1 ? seq1 : seq3;
    int c[4] = {0, 0, 0, 0};
    for (int i = 0; i < 18 && k > 0; i++) {
        const int d = seq[i];
        if (g.X[d] % (2 << c[d]) == 0) {
            c[d]++;
            k--;
        }
    }
    if (k > 0) return;
    b1 = c[1];
    b2 = c[2];
    b3 = c[3];
    spt = m;
}

struct OpsTable {
    int nc, storage, real_bytes;
    // one (parity, mu) heatbath (hb=true) or overrelaxation launch
    // sweep_dev, when not null, overrides sweep with a device-resident counter (graph replay)
    hipError_t (*update)(const Field&, const Geom&, int parity, int mu, bool hb, uint32_t sweep,
                         const uint32_t* sweep_dev, double bnc, uint64_t seed, const LaunchCfg&, hipStream_t);
    hipError_t (*cold_start)(const Field&, int nlinks, hipStream_t);
    hipError_t (*hot_start)(const Field&, const Geom&, uint64_t seed, hipStream_t);
    hipError_t (*reunitarize)(const Field&, int nlinks, hipStream_t);
    // per-block partial sums (double2) of the plaquette (spatial, temporal) and of
    // the Polyakov loop (re, im); the number of partials is *_blocks(geom)
    hipError_t (*plaquette_partial)(const Field&, const Geom&, double2* partials, hipStream_t);
    hipError_t (*polyakov_partial)(const Field&, const Geom&, double2* partials, hipStream_t);
    hipError_t (*to_canonical)(const Field&, int nlinks, double* out, hipStream_t);
    hipError_t (*from_canonical)(const Field&, int nlinks, const double* in, hipStream_t);
    // batched: every lattice of the table, one (parity, mu) launch; piped 0 plain,
    // 1 / 2 software-pipelined staple (only where has_piped); brick: the entries'
    // update bricks (ub1 >= 0) instead of the linear site order; lds: unused dynamic
    // LDS per workgroup (occupancy cap)
    hipError_t (*batch_update)(const BatchEntry* dev_table, int nlat, int total_blocks, int parity,
                               int mu, bool hb, int piped, int brick, int lds, hipStream_t);
    hipError_t (*batch_reunitarize)(const BatchEntry* dev_table, int nlat, int total_blocks, hipStream_t);
    hipError_t (*batch_plaquette_partial)(const BatchEntry* dev_table, int nlat, int total_blocks, hipStream_t);
    hipError_t (*batch_polyakov_partial)(const BatchEntry* dev_table, int nlat, int total_blocks, hipStream_t);
    // link trace / determinant sums and unitarity deviations, 3 * monitor_blocks(geom) partials
    hipError_t (*monitor_partial)(const Field&, const Geom&, double2* partials, hipStream_t);
    // plaquette partials at [0, plaquette_blocks) and Polyakov partials after them,
    // in one launch
    hipError_t (*measure_partial)(const Field&, const Geom&, double2* partials, hipStream_t);
    // measurement chain: one APE direction in -> out; multihit in -> out (18-real
    // field of the same N and precision); Wilson lines of radius r (3V-entry 18-real
    // field, r = 0 sets the identity) and the loops of radius r, tmax + 1 extents in
    // time, 3 * (tmax + 1) * monitor_blocks(geom) partials
    hipError_t (*ape)(const Field& in, const Field& out, const Geom&, int mu, double w, int nhit, double tol,
                      hipStream_t);
    // one stout direction in -> out (SU(3) variants; null otherwise)
    hipError_t (*stout)(const Field& in, const Field& out, const Geom&, int mu, double rho, hipStream_t);
    // HYP: level-1 and level-2 fields (device tables of 12 plane views of V 18-real
    // links each) from in, then the directions [mu_lo, mu_hi) of out
    hipError_t (*hyp)(const Field& in, const Field* bar, const Field* til, const Field& out, const Geom&, int mu_lo,
                      int mu_hi, double a1, double a2, double a3, int nhit, double tol, hipStream_t);
    hipError_t (*multihit)(const Field& in, const Field& out, const Geom&, uint64_t seed, int nhit, double bnc,
                           hipStream_t);
    // spatial links -> mean of nhit overrelaxation steps, temporal copied (out 18-real)
    hipError_t (*multihit_space)(const Field& in, const Field& out, const Geom&, int nhit, hipStream_t);
    // temporal links -> mean of the centre of nhit heatbath sweeps of each site's 3^3
    // box of temporal links, spatial copied (out 18-real)
    hipError_t (*multihit_ext)(const Field& in, const Field& out, const Geom&, uint64_t seed, int nhit, double bnc,
                               hipStream_t);
    hipError_t (*wl_lines)(const Field& f, const Field& lines, const Geom&, int r, hipStream_t);
    hipError_t (*wl_loops)(const Field& f, const Field& lines, const Geom&, int r, int tmax, double2* partials,
                           hipStream_t);
    // operator-basis Wilson loops: the operator fields of direction mu at radius r
    // (family 0: 10 slots, 1: 25; V 18-real entries per slot, slot k at k * stride
    // bytes after ops0 in every plane; rt8 = 2 sqrt 2), and the correlator of every
    // slot pair, nop^2 (tmax + 1) * monitor_blocks(geom) FP64 partials
    hipError_t (*wlop_fields)(const Field& f, const Field& ops0, size_t stride, const Geom&, int family, int r,
                              int mu, double rt8, hipStream_t);
    hipError_t (*wlop_corr)(const Field& f, const Field& ops0, size_t stride, const Geom&, int nop, int r, int mu,
                            int tmax, double* partials, hipStream_t);
    int max_block;   // largest workgroup the update kernels are compiled for
    int has_piped;   // the variant has the software-pipelined overrelaxation kernel
    int piped_max_block;
};

// The launch shape of batched updates unless the caller picks one (kernel 0
// heatbath, 1 overrelaxation).  Staple kernel (0 plain, 1 / 2 direction /
// half-direction pipelined): measured over 64 FP32 SU(3) lattices of 8^4..24^4 and a
// 32^4 lattice alone, half-direction pipelining is the fastest for both kernels with
// 12-real links (+3% and +17% over plain); with 8-real links for the
// overrelaxation only (+13% over 64 lattices), the pipelined heatbath being slower.
inline int default_batch_piped(const OpsTable* ops, int kernel) {
    if (!ops->has_piped) return 0;
    if (ops->storage == 12) return 2;
    return kernel == 0 ? 0 : 2;
}
// Bricks (where a 256-thread brick tiles a lattice): a 32^4 lattice runs 12% faster
// with them (its workgroups' shared neighbour links stay in one L2); over the
// 64-lattice mix they change the rate by -0.8..+1.8%.
inline int default_batch_brick(const OpsTable*, int) { return 1; }
inline int default_batch_lds(const OpsTable*, int) { return 0; }

// gfx950 dispatches workgroup b of a grid to XCD (b mod 8); each XCD has its own
// 4 MB L2.  xcd_remap turns the dispatch order into a contiguous range of tiles per
// XCD, so the lattice neighbours that share links (+-1 site, +-1 row, +-1 plane) run
// on the same L2.  A bijection of [0, nblocks) for any nblocks.
constexpr int kXcds = 8;

SUN_HD int xcd_remap(int b, int nblocks) {
    const int xcd = b % kXcds, local = b / kXcds;
    const int per = nblocks / kXcds, rem = nblocks % kXcds;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

inline int plaquette_blocks(const Geom& g) { return (g.V + kReduceBlock - 1) / kReduceBlock; }
// Polyakov loop workgroups: kReduceBlock threads = kPolySites spatial sites x
// kPolySegs segments of each time line
constexpr int kPolySegs = 4;
constexpr int kPolySites = kReduceBlock / kPolySegs;
inline int polyakov_blocks(const Geom& g) {
    int ts = g.X[0] * g.X[1] * g.X[2];
    return (ts + kPolySites - 1) / kPolySites;
}
// maximum that keeps a NaN (fmax drops it): a monitor must not report a field that
// went non-finite as an exact one
SUN_HD double nan_max(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}

// link monitor workgroups (one site per thread) and the partials every reduction of
// a lattice needs: the monitor writes three segments of monitor_blocks entries
inline int monitor_blocks(const Geom& g) { return (g.V + kReduceBlock - 1) / kReduceBlock; }
inline int partial_slots(const Geom& g) {
    const int a = 3 * monitor_blocks(g), b = plaquette_blocks(g) + polyakov_blocks(g);
    return a > b ? a : b;
}
inline int update_blocks(const Geom& g, int block) { return (g.half + block - 1) / block; }
inline int link_blocks(const Geom& g) { return (4 * g.V + kReduceBlock - 1) / kReduceBlock; }

// sum of partials[0..n) in a fixed order, times `scale`, into out (device)
hipError_t launch_finalize(const double2* partials, int n, double scale0, double scale1, double2* out,
                           hipStream_t s);
// end of a measured generation iteration, one workgroup: plaquette (partials
// [0, pb)) and Polyakov loop (partials [pb, pb + qb)) summed as launch_finalize sums,
// scaled into out[0], out[1], recorded in history row ctr - first, then ctr += 1
hipError_t launch_finalize_iteration(const double2* partials, int pb, int qb, double s_plaq, double s_poly,
                                     double2* out, double2* history, uint32_t* ctr, uint32_t first, hipStream_t s);
// Wilson loops of one radius: out[T] = scale * sum over mu < 3 and the nb partials
// of (mu, T), T = 0..tmax, in a fixed order
hipError_t launch_wl_finalize(const double2* partials, int nb, int tmax, double scale, double2* out, hipStream_t s);
// accum[o] += the sum of partials[o nb .. o nb + nb) in a fixed order, o < nout
hipError_t launch_wlop_finalize(const double* partials, int nb, int nout, double* accum, hipStream_t s);
// link monitor: out[0] = trace sum * scale_tr, out[1] = determinant sum * scale_det,
// out[2] = (max deviation, max |det - 1|) from the three segments of n partials
hipError_t launch_monitor_finalize(const double2* partials, int n, double scale_tr, double scale_det, double2* out,
                                   hipStream_t s);
// finalize for every lattice of a batch table: out[i] = (sum of entry i's partials)
// * entry i's scale, in the same order as launch_finalize
hipError_t launch_batch_finalize(const BatchEntry* table, int nlat, double2* out, hipStream_t s);
// plaquette and Polyakov finalize, history record and sweep advance of a group whose
// members' plaquette and Polyakov partials are separate buffers, in one launch
hipError_t launch_group_close(BatchEntry* t_plaq, const BatchEntry* t_poly, BatchEntry* t_update, int nlat,
                              double2* plaq, double2* poly, hipStream_t s);
// entry i's history row hist_iter <- {plaq[i], poly[i]}; hist_iter += 1
hipError_t launch_batch_record(BatchEntry* table, int nlat, const double2* plaq, const double2* poly,
                               hipStream_t s);

const OpsTable* ops_n2_f18_r4();
const OpsTable* ops_n2_f18_r8();
const OpsTable* ops_n3_f18_r4();
const OpsTable* ops_n3_f18_r8();
const OpsTable* ops_n3_f12_r4();
const OpsTable* ops_n3_f12_r8();
const OpsTable* ops_n3_f8_r4();
const OpsTable* ops_n3_f8_r8();
const OpsTable* ops_n4_f18_r4();
const OpsTable* ops_n4_f18_r8();

inline const OpsTable* find_ops(int nc, int storage, int real_bytes) {
    const OpsTable* all[] = {ops_n2_f18_r4(), ops_n2_f18_r8(), ops_n3_f18_r4(), ops_n3_f18_r8(),
                             ops_n3_f12_r4(), ops_n3_f12_r8(), ops_n3_f8_r4(),  ops_n3_f8_r8(),
                             ops_n4_f18_r4(), ops_n4_f18_r8()};
    for (const OpsTable* t : all)
        if (t->nc == nc && t->storage == storage && t->real_bytes == real_bytes) return t;
    return nullptr;
}

}  // namespace sun
