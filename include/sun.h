/*
 * sun.h -- C ABI of the MI355X SU(N) quenched ensemble-generation engine.
 *
 * Drop-in boundary for the generation path of nmrcardoso/sun (namespace CULQCD).
 * Each entry point names the reference interface it replaces (file:line in the
 * reference tree).  Differences in contract, on purpose:
 *   - every call returns an int status (0 = ok, < 0 = error) and never exits;
 *     sun_last_error() gives the message of the calling thread's last failure
 *     (the reference prints and calls exit(), comm_mpi.h:116-120);
 *   - geometry, beta and storage belong to a lattice handle, not to process
 *     globals (PARAMS/DEVPARAMS, constants.h:53-229), so lattices of different
 *     N, volume and beta coexist on one device;
 *   - the RNG is counter-based: a heatbath sweep is a pure function of
 *     (seed, sweep number), there is no RNG state array (random.h:42-101).
 *
 * Plain C types only.  A stream argument is a hipStream_t passed as void*; NULL
 * selects the context's own stream.  Host link buffers use the canonical layout
 *     double U[4][V][N][N][2]
 * (direction, even/odd site index eo(x) = parity(x)*V/2 + lex(x)/2 with
 * lex(x) = x0 + X0*(x1 + X1*(x2 + X2*x3)), row-major colour matrix, re/im).
 */
#ifndef SUN_H
#define SUN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define SUN_OK 0
#define SUN_ERR_INVALID (-1)     /* bad argument                                  */
#define SUN_ERR_HIP (-2)         /* HIP runtime error                             */
#define SUN_ERR_NOMEM (-3)       /* device or host allocation failed              */
#define SUN_ERR_UNSUPPORTED (-4) /* combination not supported (e.g. cold 8-real)  */
#define SUN_ERR_BUDGET (-5)      /* request exceeds the memory budget             */
#define SUN_ERR_STATE (-6)       /* wrong object state (e.g. unknown job id)      */

/* ---- enums mirroring include/modes.h ---------------------------------------- */
#define SUN_STORE_18 18 /* full N*N complex ("SOA")          */
#define SUN_STORE_12 12 /* SU(3) rows 0,1 ("SOA12")          */
#define SUN_STORE_8 8   /* SU(3) 8 reals ("SOA8")            */
#define SUN_PREC_F32 4
#define SUN_PREC_F64 8
#define SUN_SILENT 0    /* Verbosity, modes.h:30-35 */
#define SUN_SUMMARIZE 1
#define SUN_VERBOSE 2
#define SUN_DEBUG_VERBOSE 3
#define SUN_TUNE_NO 0   /* TuneMode, modes.h:38-41 */
#define SUN_TUNE_YES 1
#define SUN_LAYOUT_CANONICAL 0

typedef struct sun_ctx sun_ctx;
typedef struct sun_lattice sun_lattice;
typedef struct sun_batch sun_batch;
typedef struct sun_server sun_server;

typedef struct {
    int nc;          /* colours: 2, 3 or 4                                         */
    int dims[4];     /* X, Y, Z, T; each even (constants.cu:170-174)               */
    double beta;     /* Wilson coupling                                            */
    int storage;     /* SUN_STORE_*; 12 and 8 only for nc == 3                     */
    int precision;   /* SUN_PREC_F32 or SUN_PREC_F64                               */
} sun_lattice_desc;

/* ---- runtime ----------------------------------------------------------------- */
/* initCULQCD(gpuid, verbosity, tune)            include/comm_mpi.h:43, comm_mpi.cpp:92-184 */
int sun_init(int device, int verbosity, int tune, sun_ctx **out);
/* EndCULQCD(status) without exit()               include/comm_mpi.h:45, comm_mpi.cpp:186-210 */
int sun_finalize(sun_ctx *ctx);
/* errorCULQCD message                            include/comm_mpi.h:116-120 */
const char *sun_last_error(void);
/* library version string */
const char *sun_version(void);

/* ---- lattices ------------------------------------------------------------------ */
/* SETPARAMS(usetex, beta, nx, ny, nz, nt, verbose) + _gauge<Real>(type, Device, 4V, evenodd).Allocate
 *                                                include/constants.h:128, constants.cu:149-270,
 *                                                include/gaugearray.h:139-161, :603-648 */
int sun_lattice_create(sun_ctx *ctx, const sun_lattice_desc *desc, sun_lattice **out);
/* _gauge<Real>::Release                          include/gaugearray.h:657-688 */
int sun_lattice_destroy(sun_lattice *lat);
/* _gauge<Real>::Copy / Backup: device-to-device copy of the whole link field; the two
 * lattices must share N, extents, storage and precision   include/gaugearray.h:279, :282-314 */
int sun_lattice_copy(sun_lattice *dst, const sun_lattice *src, void *stream);
/* Copy of a field into a lattice of the same N and extents but any storage and
 * precision (FP64-18 -> FP32-12, ...): each link decoded as the update kernels read it,
 * widened to FP64 and encoded in dst's storage (12-real: rows 0 and 1; 8-real: the
 * reference's 8-parameter record) and precision.  Device to device through an FP64
 * staging buffer allocated for the call; synchronous.  The reference converts
 * between storage types only through its load/save paths (device_load_save.h:64-322,
 * io_gauge.cpp); here lockstep precision comparisons start both precisions from one
 * field without a host round trip. */
int sun_lattice_convert(sun_lattice *dst, const sun_lattice *src, void *stream);
/* device bytes of the link field (gaugearray.h:447-464 Bytes()) */
size_t sun_lattice_bytes(const sun_lattice_desc *desc);

/* Launch shape of a lattice's heatbath (kernel 0) or overrelaxation (kernel 1)
 * launches: the autotuner's choice, or one set by the caller.  brick[0] < 0 selects
 * the linear site order with `block` threads; otherwise a workgroup covers
 * 2^brick[0] x 2^brick[1] x 2^brick[2] rows of X0/2 sites along (x1, x2, x3).
 * lds: unused dynamic LDS per workgroup (caps workgroups per CU).  piped: the
 * software-pipelined kernels (FP32 12- and 8-real SU(3)): 1 pipelined by
 * directions (2 waves per SIMD), 2 by half-directions (3 waves).  Replaces the reference's
 * TuneParam override (include/tune.h:36-118); results never depend on it. */
typedef struct {
    int block, brick[3], lds, piped;
} sun_launch;
int sun_lattice_get_launch(const sun_lattice *lat, int kernel, sun_launch *out);
/* SUN_ERR_INVALID for a shape the lattice's kernels cannot run; marks the lattice
 * tuned, so the autotuner keeps the shape */
int sun_lattice_set_launch(sun_lattice *lat, int kernel, const sun_launch *cfg);
/* The context's launch shape for lattices of this description, as if the autotuner had
 * chosen it: lattices created later on the context with this description start with
 * it, and a job server on the context runs its large-volume jobs of this description
 * with it (otherwise the batch defaults).  Replaces an entry of the reference's tune
 * cache (tune.cpp:158-414).  SUN_ERR_INVALID for a shape the kernels cannot run on
 * that geometry. */
int sun_ctx_set_launch(sun_ctx *ctx, const sun_lattice_desc *desc, int kernel, const sun_launch *cfg);
/* _gauge<Real>::Init() -> ColdStart; refuses 8-real storage as the reference does
 *                                                gaugearray.cu:276-305, gaugearray.h:765-775 */
int sun_cold_start(sun_lattice *lat, void *stream);
/* RNG::Init(seed) + _gauge<Real>::Init(RNG) -> HotStart; links are Haar-random here
 *                                                random.cu:93-136, gaugearray.cu:353-420 */
int sun_hot_start(sun_lattice *lat, uint64_t seed, void *stream);

/* ---- updates --------------------------------------------------------------------- */
/* HeatBath<Real>(gauge, RNG).Run(stream): 8 launches, parity outer, mu inner.  The
 * RNG stream of a link is keyed by (seed, sweep, parity, mu, site).
 *                                                include/monte/monte.h:17-75, monte.cu:196-222 */
int sun_heatbath(sun_lattice *lat, uint64_t seed, uint32_t sweep, void *stream);
/* OverRelaxation<Real>(gauge).Run(stream)        include/monte/ovr.h:20, ovr.cu:182-208 */
int sun_overrelax(sun_lattice *lat, void *stream);
/* one (parity, mu) launch of each (HeatBath::apply / OverRelaxation::apply)
 *                                                monte.cu:190-194, ovr.cu:176-180 */
int sun_heatbath_sub(sun_lattice *lat, uint64_t seed, uint32_t sweep, int parity, int mu, void *stream);
int sun_overrelax_sub(sun_lattice *lat, int parity, int mu, void *stream);
/* Reunitarize<Real>(gauge).Run(stream)           include/reunitarize.h:22-60, reunitarize.cu:78-93 */
int sun_reunitarize(sun_lattice *lat, void *stream);

/* ---- observables --------------------------------------------------------------------- */
/* PlaquetteCUB<Real>(gauge).Run() -> complex{spatial, temporal}, each / (3*N*V); blocking
 *                                                include/meas/plaquette.h:120-176, plaquette_cub.cu:226-245 */
int sun_plaquette(sun_lattice *lat, void *stream, double out[2]);
/* OnePolyakovLoop<Real>(gauge).Run() -> complex, / (N*X*Y*Z); blocking
 *                                                include/meas/polyakovloop.h:15-73, polyakovloop.cu:637-768 */
int sun_polyakov(sun_lattice *lat, void *stream, double out[2]);
/* asynchronous forms: write double[2] to DEVICE memory dev_out, no host sync */
int sun_plaquette_async(sun_lattice *lat, void *stream, double *dev_out);
int sun_polyakov_async(sun_lattice *lat, void *stream, double *dev_out);

/* ---- link monitors (meas/linktrsum.h:24-90, meas/linkdetsum.h:23-100) ------------------- */
/* One pass over the field; FP64 arithmetic on the stored links (12/8-real rows
 * reconstructed as the update kernels read them), deterministic sums.
 *   out[0], out[1]  Re, Im of sum_{x,mu} Tr U_mu(x) / (4 N V)   GaugeTrace[CUB]<Real>::Run
 *                   (linktrsum.cu:130-150)
 *   out[2], out[3]  Re, Im of sum_{x,mu} det U_mu(x) / (4 V)    GaugeDet[CUB]<Real>::Run
 *                   (linkdetsum.cu:84-101)
 *   out[4]          max over links of max_ij |(U U^+)_ij - delta_ij|
 *   out[5]          max over links of |det U - 1|
 * out[4] and out[5] have no reference counterpart: they measure how far FP32 links
 * drift from SU(N) between reunitarisations. */
int sun_link_monitor(sun_lattice *lat, void *stream, double out[6]);
/* asynchronous form: writes double[6] to DEVICE memory dev_out, no host sync */
int sun_link_monitor_async(sun_lattice *lat, void *stream, double *dev_out);
/* GaugeTrace<Real>::Run (linktrsum.cu:130-150): out = mean Tr U / N */
int sun_link_trace(sun_lattice *lat, void *stream, double out[2]);
/* GaugeDet<Real>::Run (linkdetsum.cu:84-101): out = mean det U */
int sun_link_det(sun_lattice *lat, void *stream, double out[2]);

/* ---- measurement chain (smear/ape.cu, smear/multihit.cu, meas/wilsonloop.cu) ----------- */
/* The reference's static-potential chain (test.cpp:170-230) on the resident even/odd
 * field; synchronous; scratch device memory is allocated for the call. */
/* ApplyAPEinSpace / ApplyAPEinTime (ape.cu:198-264): `steps` rounds; each smears the
 * spatial directions (time_dir 0) or the temporal one (time_dir 1) from the previous
 * round's field with link <- Proj[(U + w S) / (1 + 6w)], S the staples in the spatial
 * directions; Proj = SU3project with nhit hits for N = 3 (tol unused), project_link
 * from the old link with nhit hits and tolerance tol otherwise.  In place. */
int sun_ape_smear(sun_lattice *lat, double w, int steps, int nhit, double tol, int time_dir, void *stream);
/* ApplyStoutinSpace / ApplyStoutinTime (stout.cu:687-766), SU(3) only: `steps` rounds;
 * each replaces the spatial links (time_dir 0) or the temporal ones (time_dir 1),
 * from the previous round's field, by exp(iQ) U, Q the traceless Hermitian part of
 * -i V U^+ with V = rho x the staples in the spatial directions (Morningstar-Peardon
 * exponential, FP64 arithmetic).  In place; other N: SUN_ERR_UNSUPPORTED. */
int sun_stout_smear(sun_lattice *lat, double rho, int steps, int time_dir, void *stream);
/* ApplyHYPinSpace / ApplyHYPinTime (hyp.cu:277-368, ParamHYP defaults 0.75, 0.6, 0.3):
 * `steps` rounds of hypercubic blocking of the spatial (time_dir 0) or temporal
 * (time_dir 1) links, each from the previous round's field; every level projected by
 * project_link from the unsmeared link with nhit hits and tolerance tol.  In place. */
int sun_hyp_smear(sun_lattice *lat, double alpha1, double alpha2, double alpha3, int steps, int nhit, double tol,
                  int time_dir, void *stream);
/* ApplyMultiHit (multihit.cu:291-333): out = in with every temporal link replaced by
 * the mean of itself and nhit heatbath updates in its staple (beta of `in`), RNG
 * stream (seed, site, hit); out: another lattice of the same N, extents and precision,
 * 18-real storage (the mean is not in SU(N)). */
int sun_multihit(sun_lattice *in, sun_lattice *out, uint64_t seed, int nhit, void *stream);
/* ApplyMultiHitSpace (multihitsp.cu:246-266): out = in with every spatial link replaced
 * by the mean of nhit >= 1 successive overrelaxation steps in its spatial staple (no
 * random numbers: that kernel's heatbath call is commented out); temporal links
 * copied from in (the reference leaves out's).  out as for sun_multihit. */
int sun_multihit_space(sun_lattice *in, sun_lattice *out, int nhit, void *stream);
/* ApplyMultiHitExt (multihitext.cu:395-459, LEN_SPACE 2, LEN_TIME 1): out = in with
 * every temporal link U_3(x) replaced by the mean, over nhit >= 1 checkerboard heatbath
 * sweeps of the 27 temporal links of the spatial box x + (i, j, k, 0), |i|,|j|,|k| <= 1
 * (box parity (i+j+k)&1, 0 first; every other link frozen at its input value), of the
 * box centre after each sweep (beta of `in`, RNG stream (seed, site, hit, box link));
 * spatial links copied from in.  out as for sun_multihit.  The reference requires
 * normal-order arrays and a halo-padded copy per sub-lattice pass; here one launch on
 * the even/odd field, each site's box in LDS. */
int sun_multihit_ext(sun_lattice *in, sun_lattice *out, uint64_t seed, int nhit, void *stream);
/* WilsonLoop (wilsonloop.cu:494-516): out[2 (T + r (tmax+1)) + {0,1}] = Re, Im of
 * W(r, T) = sum_{x, mu<3} Tr(L_r(x) t1 L_r(x+T e_3)^+ t0^+) / (3 N V) for r <= rmax,
 * T <= tmax (rmax at most the smallest spatial extent, tmax at most the temporal). */
int sun_wilson_loops(sun_lattice *lat, int rmax, int tmax, double *out, void *stream);
/* sun_wilson_loops without allocating: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under stream/graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results (bit for bit), but the
 * Wilson-line field and the partial sums live in a fixed 32 MiB device array (one per
 * GPU, held until the process ends) and W(r, T) goes to DEVICE memory dev_out (16-byte
 * aligned, 2 (rmax+1)(tmax+1) doubles, same layout as out), asynchronously on the
 * stream.  Needs 3V x 2N^2 x precision bytes for the line field plus 48 (tmax+1)
 * ceil(V/256) for the partials: SU(3) up to 16^4 in FP64 and 16^3x32 in FP32;
 * larger requests fail with SUN_ERR_BUDGET.  Calls on one GPU share the array: calls
 * that overlap in time on one GPU (other streams, other threads) race. */
int sun_wilson_loops_noalloc(sun_lattice *lat, int rmax, int tmax, double *dev_out, void *stream);

/* ---- operator-basis Wilson loops (src/wl/, include/meas/wlop.h, wloopex.h) ------------------
 * An operator field holds, for one spatial direction mu and radius r, nop gauge-covariant
 * sums of paths from x to x + r mu at every site (18-real links in the lattice's
 * precision): family SUN_WLOP_A0 (calcop_dg_A0.cu:46-135 with mop/mo0.cuh) has 10 slots,
 * the straight line, MO0(l) for l = 1..8 (the four staples of height l over the line,
 * / 2) and a zero slot; SUN_WLOP_33 (calcop_dg_33.cu:95-860) has 25, the line and three
 * families of staple-shaped paths combined with the reference's phase tables (slots
 * 1-8 invariant under a 90-degree rotation about mu, 9-16 times i, 17-24 times -1,
 * except slot 18, whose table entries mix the rotation phases as the reference writes
 * them).  FP64 fields equal the CPU restatement bit for bit.  The correlator adds, for
 * every slot pair (i, j) and T = 0..tmax, sum_x Re Tr(O_i(x) t1 O_j(x + T e3)^+ t0^+)
 * (t0, t1 the T temporal links from x and x + r mu) to a device accumulator
 * acc[(i nop + j)(tmax + 1) + T], in a fixed summation order. */
enum { SUN_WLOP_A0 = 0, SUN_WLOP_33 = 1 };
typedef struct sun_wlop sun_wlop;
/* Sigma_g_plus<Real> (wloopex.h:13-36; for SUN_WLOP_33 the fieldOp/wloop pair of
 * wlop.h): operator field (nop V links) and accumulator (nop^2 (tmax+1) doubles, zero)
 * on the lattice's device; tmax <= the temporal extent */
int sun_wlop_create(sun_lattice *lat, int family, int tmax, sun_wlop **out);
int sun_wlop_destroy(sun_wlop *h);
int sun_wlop_nops(const sun_wlop *h);
/* CalcWLOPs_A0 (calcop_dg_A0.cu:307-343) / CalcWLOPs_dg_33 (calcop_dg_33.cu:966-987):
 * the operator field of direction mu (0..2) at radius r (1..X_mu), asynchronous */
int sun_wlop_fields(sun_wlop *h, int r, int mu, void *stream);
/* CalcWilsonLoop_A0 (wilsonloop_dg_A0.cu:57-105, 219-254) / CalcWilsonLoop_dg
 * (wilsonloop_dg.cu:82-129, 227-266): adds the correlators of the current operator
 * field (t1 at distance r along mu) to the accumulator, asynchronous */
int sun_wlop_correlate(sun_wlop *h, int r, int mu, void *stream);
/* cudaMemset(arg.wloop, 0) of test.cpp:280 */
int sun_wlop_clear(sun_wlop *h, void *stream);
/* the accumulator, nop^2 (tmax + 1) doubles (synchronous) */
int sun_wlop_read(sun_wlop *h, double *out, void *stream);
/* the accumulator in the component-table order of wilsonloop_dg.cu:101-122
 * (DEV_Ops / DEV_OpComps / DEV_OpPos, which no reference file sets; here arguments):
 * for component c, pairs (pos[c] + a, pos[c] + b), a, b < comps[c], a outer;
 * out[id (tmax + 1) + T] with id counting those pairs */
int sun_wlop_read_blocks(sun_wlop *h, int ncomp, const int *comps, const int *pos, double *out, void *stream);
/* the operator field in the canonical FP64 order: out[((k V) + s) N^2 2], s the
 * even/odd site (synchronous) */
int sun_wlop_fields_to_host(sun_wlop *h, double *out);
/* test.cpp:273-298: for r = 2..rmax, the accumulator after clearing it and adding
 * mu = 0, 1, 2, divided by 3 N V: out[(((r - 2) nop^2 + i nop + j)(tmax + 1)) + T] */
int sun_wilson_loops_ops(sun_lattice *lat, int family, int rmax, int tmax, double *out, void *stream);

/* ---- host exchange (gaugearray.h:521-596 Get/Set, canonical FP64 layout) ---------------- */
int sun_links_to_host(sun_lattice *lat, double *host, int layout);
int sun_links_from_host(sun_lattice *lat, const double *host, int layout);

/* ---- checkpoints ------------------------------------------------------------------------
 * SUN_FILE_NATIVE: the lattice's device bytes exactly, an 80-byte header (N, extents,
 * storage, precision, beta, seed, next sweep number, CRC32 of the payload); loading
 * restores the field bit for bit and returns (seed, next sweep), which is the whole RNG
 * state, so a chain resumes exactly.  SUN_FILE_REFERENCE[_NOHEADER]: the layout of the
 * reference's SaveBin_Gauge / ReadBin_Gauge (io_gauge.cpp:429-583, :605-700): [header
 * grid[4], beta, bytes per real] then every site in lexicographic order (x fastest) with
 * its four full N x N complex matrices as float (save_prec 4) or double (8); reading
 * checks grid and beta (io_gauge.cpp:585-593) and projects onto the lattice's storage.
 * Both synchronise the device first.  seed / next_sweep are ignored (save) or left
 * unchanged (load) for the reference formats; save_prec is read from the header when
 * there is one. */
#define SUN_FILE_NATIVE 0
#define SUN_FILE_REFERENCE 1
#define SUN_FILE_REFERENCE_NOHEADER 2
int sun_lattice_save(sun_lattice *lat, const char *path, int format, int save_prec, uint64_t seed,
                     uint32_t next_sweep);
int sun_lattice_load(sun_lattice *lat, const char *path, int format, int save_prec, uint64_t *seed,
                     uint32_t *next_sweep);

/* ---- generation driver ------------------------------------------------------------------
 * The loop the reference's API implies but does not ship (no HB/OR driver exists in
 * test.cpp): for it in [0, n_iter): heatbath (sweep = first_sweep + it), n_or
 * overrelaxation sweeps, reunitarize every reunit_every iterations (0 = never),
 * plaquette and Polyakov loop every iteration.  obs (host, may be NULL) receives
 * n_iter rows of {plaq_spatial, plaq_temporal, poly_re, poly_im}. */
typedef struct {
    uint64_t seed;
    uint32_t first_sweep;
    int n_iter;
    int n_or;
    int reunit_every;
    int measure;     /* 0: no observables, 1: plaquette + Polyakov each iteration */
    int use_graph;   /* capture one iteration in a hipGraph and replay it           */
} sun_generate_desc;
int sun_generate(sun_lattice *lat, const sun_generate_desc *desc, void *stream, double *obs);

/* ---- batched lattices: one launch per (parity, mu) covers every lattice -------------------
 * All lattices of a batch share N, storage and precision; volumes and beta differ. */
int sun_batch_create(sun_ctx *ctx, int n, sun_lattice *const *lats, const uint64_t *seeds, sun_batch **out);
int sun_batch_destroy(sun_batch *b);
/* launch shape of a batch's heatbath (kernel 0) or overrelaxation (kernel 1)
 * launches.  piped: the staple kernel, 0 plain, 1 / 2 direction / half-direction
 * software-pipelined (FP32 SU(3) 12- and 8-real only; SUN_ERR_INVALID otherwise).
 * brick: 1 maps each 256-thread workgroup to a brick of rows of its lattice where one
 * tiles it (X0/2 a power of two), 0 the linear site order.  lds: unused dynamic LDS
 * per workgroup, 0..65536 bytes, capping the workgroups resident per CU.  Every
 * shape gives bit-identical links.  Replaces the block-size search of
 * tune.cpp:158-414 over the sweep launches of monte.cu:224-226 */
int sun_batch_get_launch(const sun_batch *b, int kernel, int *piped, int *brick, int *lds);
int sun_batch_set_launch(sun_batch *b, int kernel, int piped, int brick, int lds);
int sun_batch_heatbath(sun_batch *b, const uint32_t *sweeps, void *stream);
int sun_batch_overrelax(sun_batch *b, void *stream);
int sun_batch_reunitarize(sun_batch *b, void *stream);
/* out: double[2*n] plaquette (spatial, temporal) per lattice, blocking */
int sun_batch_plaquette(sun_batch *b, void *stream, double *out);
/* n_iter x (HB + n_or OR [+ reunit]); hipGraph-captured when use_graph; per-lattice
 * plaquette of every iteration into obs (host, n_iter*2*n doubles, may be NULL).
 * Chunks of lattices run on the caller's stream and two streams of the batch, which
 * wait for the caller's stream first and which it waits for at the end. */
int sun_batch_generate(sun_batch *b, uint32_t first_sweep, int n_iter, int n_or, int reunit_every,
                       int use_graph, void *stream, double *obs);

/* ---- job server: continuous batching of ensemble-generation jobs ------------------------
 * Jobs of different N, volume, beta and length are admitted while their link field
 * and scratch fit the HBM budget (reserved once at creation; no device allocation
 * afterwards).  Every scheduling step runs one iteration (1 heatbath + n_or
 * overrelaxation sweeps, reunitarisation when reunit_every == 1, plaquette and
 * Polyakov loop) of every resident job, jobs of one kind (N, storage, precision,
 * n_or) together in one launch per (parity, mu), kinds on separate streams.  A job
 * finishing frees its memory for the next pending one at the step boundary.  A
 * job's results do not depend on what else is resident (bit-identical alone and in
 * the mix).  No counterpart in the reference, which runs one lattice per process. */
typedef struct {
    int nc, dims[4];
    double beta;
    int storage, precision;
    uint64_t seed;
    int hot_start;      /* 1: Haar start from seed; 0: cold start                    */
    int n_iter;         /* iterations (>= 1)                                          */
    int n_or;           /* overrelaxation sweeps per iteration                        */
    int reunit_every;   /* 0 or 1                                                      */
} sun_job_desc;

#define SUN_JOB_PENDING 0
#define SUN_JOB_RUNNING 1
#define SUN_JOB_DONE 2
#define SUN_JOB_CANCELLED 3
#define SUN_JOB_SUSPENDED 4
#define SUN_JOB_FAILED 5    /* observables not finite or out of range, or its suspend file
                               could not be written: sun_server_job_error says which */

typedef struct {
    uint64_t id;
    int state, iters_done;
    double t_submit, t_start, t_end;   /* seconds on the server clock                   */
    double plaq[2], poly[2];           /* last iteration (done jobs)                     */
} sun_job_status;

typedef struct {
    uint64_t jobs_done, link_updates, steps;
    double busy_s;                     /* wall time spent inside scheduling steps        */
    size_t budget, bytes_in_use;
    int running, pending;
} sun_server_counters;

int sun_server_create(sun_ctx *ctx, size_t budget_bytes, int max_jobs, sun_server **out);
/* Checks a job description as sun_server_submit does, without a server or a device,
 * and gives the server memory the job occupies while resident (link field, reduction
 * scratch, observable history; arena-rounded): for sizing a budget.  Same status
 * codes as submit, except SUN_ERR_BUDGET. */
int sun_job_bytes(const sun_job_desc *job, size_t *bytes);
int sun_server_destroy(sun_server *srv);
/* SUN_ERR_BUDGET if the job alone exceeds the budget; SUN_ERR_UNSUPPORTED for a cold
 * 8-real start, reunit_every > 1, or 12-real links with reunit_every = 0 (they leave
 * SU(3) geometrically and overflow within about a dozen iterations) */
int sun_server_submit(sun_server *srv, const sun_job_desc *job, uint64_t *id);
int sun_server_cancel(sun_server *srv, uint64_t id);
int sun_server_step(sun_server *srv);                 /* one step on the calling thread  */
int sun_server_start(sun_server *srv);                /* background scheduling thread    */
int sun_server_stop(sun_server *srv);
int sun_server_wait(sun_server *srv, uint64_t id, double timeout_s);  /* SUN_ERR_STATE on timeout */
int sun_server_status(sun_server *srv, uint64_t id, sun_job_status *st);
/* rows {plaq_spatial, plaq_temporal, poly_re, poly_im} of every iteration of a done job */
int sun_server_observables(sun_server *srv, uint64_t id, double *out, int max_rows);
int sun_server_release(sun_server *srv, uint64_t id);  /* forget a finished job          */
/* Why a job is in state SUN_JOB_FAILED (empty otherwise): its latest observables were
 * not finite or outside [-1, 1] (checked after every group launch and over the whole
 * history when it completes), or its suspend file could not be written.  The
 * reference stops the process on such errors (errorCULQCD, comm_mpi.h:116-120); the
 * server fails the one job and keeps running the others. */
int sun_server_job_error(sun_server *srv, uint64_t id, char *buf, int len);
/* Takes a job off the device into a file: a pending job at once, a running one at its
 * next iteration boundary (wait for state SUN_JOB_SUSPENDED; the path must be creatable
 * when this is called, else SUN_ERR_INVALID and the job runs on).  The file holds the job
 * description, the iterations done, their observable rows and a native checkpoint of
 * the lattice (SUN_FILE_NATIVE layout).  No counterpart in the reference; its
 * checkpoint format is io_gauge.cpp:429-700. */
int sun_server_suspend(sun_server *srv, uint64_t id, const char *path);
/* Submits a suspended job again, on this or another server: it continues from where it
 * stopped and its results equal those of the job never suspended. */
int sun_server_resume(sun_server *srv, const char *path, uint64_t *id);
int sun_server_stats(sun_server *srv, sun_server_counters *st);
double sun_server_clock(sun_server *srv);

/* ---- streams and events (timing helpers for callers without a HIP binding) ---------- */
int sun_stream_create(void **stream);
int sun_stream_destroy(void *stream);
int sun_stream_synchronize(void *stream);
int sun_event_create(void **ev);
int sun_event_destroy(void *ev);
int sun_event_record(void *ev, void *stream);
int sun_event_elapsed_ms(void *start, void *stop, float *ms);
int sun_device_name(sun_ctx *ctx, char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* SUN_H */
