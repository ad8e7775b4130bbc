/*
 * sun_oracle.h -- FP64 CPU restatement of the quenched SU(N) ensemble-generation
 * path of nmrcardoso/sun (CULQCD).  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP engine is compared against.  It is linked by the tests, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.
 *
 * Host link layout used throughout ("canonical layout"):
 *     double U[4][V][N][N][2]      (mu-major, then even/odd site index, row-major
 *                                    complex entries, re/im interleaved)
 * where the even/odd site index of x is  eo(x) = parity(x)*V/2 + lex(x)/2,
 * lex(x) = x0 + N0*(x1 + N1*(x2 + N2*x3)),  as in include/index.h:133-141,171-183.
 *
 * Every function states the reference file:line it restates.  Pinning of each
 * function (known-answer vectors, closed forms, invariants, brute force) lives in
 * tests/test_oracle_*.py.
 */
#ifndef SUN_ORACLE_H
#define SUN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Storage codes, matching the reference's ArrayType (include/modes.h): full
 * N*N complex ("SOA", 18 reals for SU(3)), 12-real ("SOA12") and 8-real ("SOA8"). */
enum { OR_STORE_18 = 18, OR_STORE_12 = 12, OR_STORE_8 = 8 };

/* ---- counter-based RNG (Philox4x32-10, Salmon et al. SC'11) ---------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* uniform double in (0,1] from one 32-bit Philox word: (w + 1) * 2^-32 */
double or_u01(uint32_t w);
/* n uniforms of the per-site stream (seed, c0, c2, c3), draw counter starting at 0 */
void or_stream_uniforms(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double *out);

/* ---- SU(2) sampler (device_PHB_OVR.h:162-223), driven by a uniform list ----- */
/* Draws from u[] sequentially; returns number of uniforms consumed; a[4] out.    */
int or_generate_su2_milc(double al, const double *u, int nu, double a[4]);
/* n samples of a0 at weight al, drawn from stream (seed, i, 0, 0x30000000). */
void or_sample_a0(double al, uint64_t seed, int n, double *a0_out);

/* ---- geometry (include/index.h:133-202) ----------------------------------- */
void or_eo_to_coords(int id, int parity, const int dims[4], int x[4]);
int  or_coords_to_eo(const int x[4], const int dims[4]);
int  or_neighbor_eo(int id, int parity, const int dims[4], int mu, int lmu, int nu, int lnu);

/* ---- field initialisation -------------------------------------------------- */
void or_cold_start(int nc, const int dims[4], double *U);             /* gaugearray.cu:276-305 */
void or_hot_start(int nc, const int dims[4], uint64_t seed, int storage, double *U); /* Haar */

/* ---- storage codecs (device_load_save.h:76-322, reconstruct_12p_8p.h:17-82) - */
void or_encode12(const double *link, double *p12);       /* rows 0,1: 12 reals        */
void or_decode12(const double *p12, double *link);       /* row 2 = conj(r0 x r1)      */
void or_encode8(const double *link, double *p8);         /* e01,e02,e10,(arg e00,arg e20) */
void or_decode8(const double *p8, double *link);
void or_project_storage(int nc, int storage, long nlinks, double *U);  /* decode(encode(U)) */

/* ---- staple (monte.cu:63-80) ------------------------------------------------ */
void or_staple(int nc, const int dims[4], const double *U, int parity, int mu, int id, double *S);

/* ---- one (parity,mu) sub-sweep and full sweeps ------------------------------ */
/* heatbath: monte.cu:44-88 + device_PHB_OVR.h:361-485; overrelax: ovr.cu:39-80 +
 * device_PHB_OVR.h:493-608.  Returns 0, or <0 on invalid arguments.             */
int or_heatbath_sub(int nc, const int dims[4], double beta, uint64_t seed, uint32_t sweep,
                    int storage, double *U, int parity, int mu);
int or_overrelax_sub(int nc, const int dims[4], int storage, double *U, int parity, int mu);
/* parity-outer, mu-inner schedule (monte.cu:205-206, ovr.cu:191-192) */
int or_heatbath_sweep(int nc, const int dims[4], double beta, uint64_t seed, uint32_t sweep,
                      int storage, double *U);
int or_overrelax_sweep(int nc, const int dims[4], int storage, double *U);
/* the heatbath sweep, also writing margins[4V]: for every updated link the smallest
 * distance from a tie among its sampler's decisions (branch al > 2, accept tests);
 * +inf for links the sweep does not update.  For attributing FP32 accept/reject flips */
int or_heatbath_sweep_margins(int nc, const int dims[4], double beta, uint64_t seed, uint32_t sweep,
                              int storage, double *U, double *margins);

/* ---- single-link updates (exposed for closed-form tests) --------------------- */
void or_overrelax_link(int nc, double *U, const double *F);
/* heatbath of one link; uniforms from stream (seed, c0, c2, c3) */
void or_heatbath_link(int nc, double beta, double *U, const double *F,
                      uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3);

/* ---- reunitarisation (reunitlink.h:16-85; SU(2) textbook projection) -------- */
void or_reunit_link(int nc, double *link);
void or_reunitarize(int nc, long nlinks, int storage, double *U);

/* ---- observables ------------------------------------------------------------ */
/* plaquette_cub.cu:96-149,234: out[0]=spatial, out[1]=temporal, each / (3*N*V) */
void or_plaquette(int nc, const int dims[4], const double *U, double out[2]);
/* polyakovloop.cu:53-99,651: complex mean trace / (N*Nx*Ny*Nz) */
void or_polyakov(int nc, const int dims[4], const double *U, double out[2]);

/* linktrsum.cu:130-150, linkdetsum.cu:84-101: out = {mean Tr U / N (re, im),
 * mean det U (re, im), max_links max_ij |(U U^+ - 1)_ij|, max_links |det U - 1|};
 * U holds nlinks = 4V links in the canonical layout */
void or_link_monitor(int nc, long V, const double *U, double out[6]);

/* ---- measurement chain (smear/ape.cu, smear/multihit.cu, meas/wilsonloop.cu) -- */
/* projectlink.h:16-64: w <- SU(N) element maximising Re Tr(w q^+) by nhit rounds of
 * SU(2)-subgroup hits from w (tol > 0: early stop, the reference's measure) */
void or_project_link(int nc, double *w, const double *q, int nhit, double tol);
/* projectlink.h:119-139: F <- its SU(3) projection (hits from the identity) */
void or_su3_project(double *F, int nhit);
/* ape.cu:198-264: steps of APE smearing, spatial (time_dir 0: mu = 0,1,2 with
 * spatial staples) or temporal (time_dir 1: mu = 3), weight w; 18-real U in place */
int or_ape_smear(int nc, const int dims[4], double *U, double w, int steps, int nhit, double tol, int time_dir);
/* exp(iQ), Q traceless Hermitian 3x3 (Morningstar-Peardon; stout.cu:94-466) */
void or_exp_iq(const double *Q, double *E);
/* stout.cu:687-766: steps of stout smearing with weight rho, spatial (time_dir 0) or
 * temporal (time_dir 1) links, SU(3) 18-real U in place */
int or_stout_smear(const int dims[4], double *U, double rho, int steps, int time_dir);
/* hyp.cu:277-368: steps of HYP smearing (alphas a1, a2, a3; project_link from the
 * unsmeared link with nhit hits, tolerance tol) of the spatial (time_dir 0) or
 * temporal (time_dir 1) links; 18-real U in place.  The reference's recursion. */
int or_hyp_smear(int nc, const int dims[4], double *U, double a1, double a2, double a3, int steps, int nhit,
                 double tol, int time_dir);
/* multihit.cu:112-147: Uout = Uin with each temporal link replaced by the mean of
 * itself and nhit heatbath updates in its staple (stream (seed, id, hit, TAG_MH|p<<4|3)) */
int or_multihit(int nc, const int dims[4], double beta, uint64_t seed, int nhit, const double *Uin, double *Uout);
/* the same, also writing margins[4V] (temporal links: the smallest distance from a tie
 * among the sampler decisions of its nhit updates; other entries untouched) */
int or_multihit_margins(int nc, const int dims[4], double beta, uint64_t seed, int nhit, const double *Uin,
                        double *Uout, double *margins);
/* multihitsp.cu:115-161: Uout = Uin with every spatial link replaced by the mean of
 * nhit successive overrelaxation steps in its spatial staple */
int or_multihit_space(int nc, const int dims[4], int nhit, const double *Uin, double *Uout);
/* multihitext.cu:241-330, :395-459: Uout = Uin with each temporal link replaced by the
 * mean, over nhit checkerboard heatbath sweeps of the 27 temporal links of its 3^3
 * spatial box (other links frozen at their input values), of the box centre after
 * each sweep (streams (seed, idd, hit, TAG_MHX | box index)) */
int or_multihit_ext(int nc, const int dims[4], double beta, uint64_t seed, int nhit, const double *Uin, double *Uout);
/* the same, margins[4V] as for or_multihit_margins over every update in the link's box */
int or_multihit_ext_margins(int nc, const int dims[4], double beta, uint64_t seed, int nhit, const double *Uin,
                            double *Uout, double *margins);
/* wilsonloop.cu:44-121,494-516: out[2(T + r(tmax+1)) + {0,1}] = W(r, T), r <= rmax,
 * T <= tmax, averaged over sites and the three spatial directions, / (3 N V) */
int or_wilson_loops(int nc, const int dims[4], const double *U, int rmax, int tmax, double *out);

/* ---- operator-basis Wilson loops (src/wl/calcop_dg_A0.cu, src/wl/calcop_dg_33.cu,
 * include/mop/mo0.cuh, src/wl/wilsonloop_dg_A0.cu, src/wl/wilsonloop_dg.cu) ---- */
enum { OR_WLOP_A0 = 0, OR_WLOP_33 = 1 };
/* the operator fields of spatial direction mu (0..2) at radius r >= 1:
 * ops[((slot V) + s) N^2 2], s the even/odd site; family OR_WLOP_A0 has 10 slots
 * (calcop_dg_A0.cu:46-135), OR_WLOP_33 25 (calcop_dg_33.cu:95-860) */
int or_wlop_fields(int nc, const int dims[4], const double *U, int family, int r, int mu, double *ops);
/* wilsonloop_dg_A0.cu:57-105 over every slot pair: out[(i nop + j)(tmax + 1) + T] +=
 * sum over sites x of Re Tr(O_i(x) t1 O_j(x + T e3)^+ t0^+), t0 (t1) the product of T
 * temporal links from x (x + r mu) */
int or_wlop_correlate(int nc, const int dims[4], const double *U, const double *ops, int nop, int r, int mu,
                      int tmax, double *out);

/* ---- helpers ---------------------------------------------------------------- */
void or_matmul(int nc, const double *A, const double *B, double *C);     /* C = A*B */
void or_det(int nc, const double *A, double out[2]);
int  or_num_threads(void);
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
