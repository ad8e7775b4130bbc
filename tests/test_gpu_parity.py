"""GPU parity: the HIP engine, called through the C ABI, against the FP64 oracle.

Tolerances: link-by-link after one overrelaxation or heatbath sweep
<= 1e-12 in FP64 storage and <= 5e-5 in FP32 storage, on thermalised fields (the
operating regime: oracle cold start + 12 heatbath/overrelaxation iterations).

Haar-random fields are ill-conditioned for these updates: the SU(2) projections of
random staples can be tiny and conj_normalize divides by them, so the oracle itself
moves by ~1e-11 (SU(3)) to ~1e-8 (SU(4)) after one sweep when its input is perturbed
by 1e-15.  In FP64 that does not matter: the device runs the oracle's operations in
the oracle's order, with the same portable transcendentals (sun_libm.h), so FP64
Haar-start sweeps are held to 1e-12 too.  FP32 deviations on Haar inputs are bounded
by a multiple of the oracle's own sensitivity to one-ulp input perturbations
(`self_sensitivity`).

In FP32 a heatbath accept/reject branch may flip relative to the FP64 oracle (the
sampler runs in FP32); a flip changes one link by O(1) and, through the staples, a
few neighbours updated after it.  FP32 heatbath tests therefore require every link
within the bar except those a flip explains: the oracle reports, per link, how close
its sampler's decisions came to a tie, and an outlier must have a near-tie decision
of its own or an outlier among the staple links updated before it (flip_report).
"""
import numpy as np
import pytest

import sun_amd

pytestmark = pytest.mark.gpu

VARIANTS = [  # (nc, storage, precision)
    (3, 18, 8), (3, 12, 8), (3, 8, 8), (2, 18, 8), (4, 18, 8),
    (3, 18, 4), (3, 12, 4), (3, 8, 4), (2, 18, 4), (4, 18, 4),
]
IDS = [f"su{n}_s{s}_{'f64' if p == 8 else 'f32'}" for n, s, p in VARIANTS]


def tol(prec):
    return 1e-12 if prec == 8 else 5e-5


def store_tol(storage, prec):
    """Rounding of one stored link: FP64 exact to ~1e-13; FP32 to a few ulp, more
    for 8-real storage, whose decode divides by |e01|^2+|e02|^2."""
    if prec == 8:
        return 1e-13
    return 2e-5 if storage == 8 else 2e-6


def link_bound(lat, U, ref, storage, prec, sweep):
    """Per-link bar for a GPU sweep from input U against the oracle's result ref.
    FP64 and 12/18-real FP32: tol(prec).  8-real FP32: a sweep projects every
    updated link to 8 reals, whose decode divides by |e01|^2+|e02|^2, so at some
    links the exact result moves by up to ~1e-4 when the input moves by one FP32 ulp
    (also for fields thermalised from a cold start).  There the bar is 5e-5 plus four
    times the FP32 storage round-trip error of ref at the link plus ten times the
    oracle's own change at the link under a one-ulp (relative 2^-23) input change."""
    base = tol(prec)
    if not (storage == 8 and prec == 4):
        return base
    lat.from_host(ref)
    store = per_link_diff(lat.to_host(), ref)
    rng = np.random.default_rng(321)
    P = U * (1 + 2.0 ** -23 * rng.standard_normal(U.shape))
    sweep(P)
    return base + 4.0 * store + 10.0 * per_link_diff(P, ref)


def maxdiff(A, B):
    return float(np.abs(A - B).max())


def per_link_diff(A, B):
    return np.abs(A - B).reshape(A.shape[0] * A.shape[1], -1).max(axis=1)


def make(ctx, nc, dims, beta, storage, prec):
    return sun_amd.Lattice(ctx, nc, dims, beta, storage, prec)


def start_config(oracle, lat, seed):
    """Hot-start the GPU lattice and return its decoded links (the common input)."""
    lat.hot_start(seed)
    return lat.to_host()


BETAS = {2: 2.3, 3: 6.0, 4: 11.0}
# FP32 heatbath links beyond the bar in one sweep of these small lattices (every one
# attributed to an accept/reject flip): 0 measured on thermalised inputs, a small margin above
MAX_FLIP_OUTLIERS = 4
_THERM = {}


def thermalised(oracle, nc, dims, storage):
    """oracle FP64 field after a cold (Haar for 8-real) start and 12 iterations"""
    key = (nc, dims, storage)
    if key not in _THERM:
        beta = BETAS[nc]
        U = oracle.hot_start(nc, dims, 77, storage) if storage == 8 else oracle.cold_start(nc, dims)
        for it in range(12):
            oracle.heatbath_sweep(U, dims, beta, 78, it, storage)
            for _ in range(2):
                oracle.overrelax_sweep(U, dims, storage)
            oracle.reunitarize(U, storage)
        _THERM[key] = U
    return _THERM[key].copy()


# A sampler decision whose FP64 distance from a tie is below this may go the other
# way in FP32 arithmetic (FP32 transcendentals err by ~1e-6 relative, the staple and
# alpha by FP32 rounding); FP32 heatbath links that differ by O(1) must come from one.
TIE = 1e-4


def staple_links(dims, parity, mu):
    """link ids mu'*V + eo of the 18 staple links of every site of the (parity, mu)
    sub-sweep, shape (18, V/2): U_nu(x), U_mu(x+nu), U_nu(x+mu), U_nu(x-nu),
    U_mu(x-nu), U_nu(x-nu+mu) for nu != mu (monte.cu:63-80)"""
    X = np.array(dims)
    V = int(np.prod(X))
    half = V // 2
    lex = np.arange(V)
    x = np.stack([lex % X[0], (lex // X[0]) % X[1], (lex // (X[0] * X[1])) % X[2], lex // (X[0] * X[1] * X[2])])
    eo_of_lex = (x.sum(axis=0) & 1) * half + lex // 2
    lex_of_eo = np.empty(V, dtype=np.int64)
    lex_of_eo[eo_of_lex] = lex
    xs = x[:, lex_of_eo[parity * half:(parity + 1) * half]]

    def eo(y):
        y = y % X[:, None]
        return eo_of_lex[y[0] + X[0] * (y[1] + X[1] * (y[2] + X[2] * y[3]))]

    e = np.eye(4, dtype=np.int64)[:, :, None]
    out = []
    for nu in range(4):
        if nu == mu:
            continue
        out += [nu * V + eo(xs), mu * V + eo(xs + e[nu]), nu * V + eo(xs + e[mu]), nu * V + eo(xs - e[nu]),
                mu * V + eo(xs - e[nu]), nu * V + eo(xs - e[nu] + e[mu])]
    return np.array(out)


def flip_report(d, bar, margins, dims, propagate=True):
    """Attribution of the FP32 links of a heatbath sweep (or a multihit) that miss
    the per-link bar.  d, bar: per link (4V,); margins: the oracle's per-link
    smallest distance from a sampler tie (shape (4, V)).  An outlier is explained by
    its own near-tie decision (margin < TIE), or -- with propagate, for sweeps -- by
    an outlier among its staple links that the sweep updated before it (parity
    outer, mu inner: the O(1) change reaches it through the staple).  Returns
    (outliers, near-tie outliers, unexplained link ids)."""
    V = int(np.prod(dims))
    half = V // 2
    out = d > bar
    tie = margins.reshape(-1) < TIE
    ids_all = np.arange(4 * V)
    order = (ids_all % V >= half) * 4 + ids_all // V
    unexplained = []
    for parity in (0, 1):
        for mu in range(4):
            ids = mu * V + parity * half + np.arange(half)
            o = out[ids]
            if not o.any():
                continue
            ok = tie[ids[o]]
            if propagate:
                nb = staple_links(dims, parity, mu)[:, o]
                ok = ok | (out[nb] & (order[nb] < parity * 4 + mu)).any(axis=0)
            unexplained += [int(i) for i in ids[o][~ok]]
    return int(out.sum()), int((out & tie).sum()), unexplained


def load_config(lat, U):
    """put U on the GPU and return the decoded stored field (the common input)"""
    lat.from_host(U)
    return lat.to_host()


def self_sensitivity(fn, U, eps):
    """max change of fn's output when its input is perturbed by relative eps"""
    rng = np.random.default_rng(123)
    A = U.copy()
    B = U * (1 + eps * rng.standard_normal(U.shape))
    fn(A)
    fn(B)
    return float(np.abs(A - B).max())


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_cold_start_observables(ctx, nc, storage, prec):
    dims = (4, 4, 4, 4)
    with make(ctx, nc, dims, 6.0, storage, prec) as lat:
        if storage == 8:
            with pytest.raises(sun_amd.SunError, match="8-real"):
                lat.cold_start()
            return
        lat.cold_start()
        p = lat.plaquette()
        assert p[0] == 1.0 and p[1] == 1.0
        assert lat.polyakov() == 1.0 + 0.0j


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_hot_start_matches_oracle(ctx, oracle, nc, storage, prec):
    dims = (4, 4, 4, 6)
    with make(ctx, nc, dims, 6.0, storage, prec) as lat:
        lat.hot_start(11)
        G = lat.to_host()
    O = oracle.hot_start(nc, dims, 11, storage)
    assert maxdiff(G, O) <= store_tol(storage, prec)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_host_roundtrip(ctx, oracle, nc, storage, prec):
    dims = (4, 4, 4, 4)
    O = oracle.hot_start(nc, dims, 5, storage)
    with make(ctx, nc, dims, 6.0, storage, prec) as lat:
        lat.from_host(O)
        G = lat.to_host()
    assert maxdiff(G, O) <= store_tol(storage, prec)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (4, 6, 8, 10)])
def test_overrelax_sweep_parity(ctx, oracle, nc, storage, prec, dims):
    with make(ctx, nc, dims, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, dims, storage))
        lat.overrelax()
        G = lat.to_host()
        ref = U.copy()
        oracle.overrelax_sweep(ref, dims, storage)
        bound = link_bound(lat, U, ref, storage, prec, lambda A: oracle.overrelax_sweep(A, dims, storage))
        assert np.all(per_link_diff(G, ref) <= bound)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (4, 6, 8, 10)])
def test_heatbath_sweep_parity(ctx, oracle, nc, storage, prec, dims):
    """FP64: 1e-12 per link.  FP32: every link within the bar (5e-5; 8-real with its
    conditioning term) except those an accept/reject flip explains (flip_report);
    the outlier counts are printed and bounded."""
    beta = BETAS[nc]
    seed, sweep = 1234, 7
    with make(ctx, nc, dims, beta, storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, dims, storage))
        lat.heatbath(seed, sweep)
        G = lat.to_host()
        ref = U.copy()
        margins = oracle.heatbath_sweep_margins(ref, dims, beta, seed, sweep, storage)
        d = per_link_diff(G, ref)
        if prec == 8:
            assert d.max() <= 1e-12
            return
        bar = link_bound(lat, U, ref, storage, prec, lambda A: oracle.heatbath_sweep(A, dims, beta, seed, sweep,
                                                                                     storage))
    n_out, n_tie, unexplained = flip_report(d, bar, margins, dims)
    print(f"heatbath FP32 su{nc} s{storage} {dims}: {n_out} of {d.size} links beyond the bar, {n_tie} at "
          f"near-tie decisions, {len(unexplained)} unexplained")
    assert not unexplained, unexplained[:10]
    assert n_out <= MAX_FLIP_OUTLIERS, n_out


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_overrelax_haar_start(ctx, oracle, nc, storage, prec):
    """From a Haar start.  FP64: the device and the oracle run the same operations
    (no contraction; the 8-real codec's atan2, sin and cos are the shared portable
    ones), so one sweep agrees to 1e-12 whatever the conditioning.  FP32: within 50x
    the oracle's own change under a one-ulp (2^-23) input perturbation."""
    dims = (4, 6, 8, 10)
    with make(ctx, nc, dims, BETAS[nc], storage, prec) as lat:
        U = start_config(oracle, lat, 21)
        lat.overrelax()
        G = lat.to_host()
    if prec == 8:
        oracle.overrelax_sweep(U, dims, storage)
        assert maxdiff(G, U) <= 1e-12
        return
    sens = self_sensitivity(lambda A: oracle.overrelax_sweep(A, dims, storage), U, 2.0 ** -23)
    oracle.overrelax_sweep(U, dims, storage)
    assert maxdiff(G, U) <= max(tol(prec), 50 * sens)


@pytest.mark.parametrize("nc,storage", [(3, 18), (3, 12), (3, 8), (2, 18), (4, 18)])
def test_heatbath_haar_start_f64(ctx, oracle, nc, storage):
    """FP64 heatbath from a Haar start: the sampler's log, exp and sin/cos are the
    portable ones the oracle uses, so every accept/reject decision and every link
    agree (1e-12) although a Haar field is ill-conditioned for the update."""
    dims = (4, 6, 8, 10)
    beta, seed = BETAS[nc], 99
    with make(ctx, nc, dims, beta, storage, 8) as lat:
        U = start_config(oracle, lat, 22)
        lat.heatbath(seed, 0)
        G = lat.to_host()
    oracle.heatbath_sweep(U, dims, beta, seed, 0, storage)
    assert maxdiff(G, U) <= 1e-12


@pytest.mark.parametrize("nc,storage,prec", [(3, 18, 8), (3, 12, 4), (2, 18, 8), (4, 18, 8)],
                         ids=["su3_s18_f64", "su3_s12_f32", "su2_f64", "su4_f64"])
def test_sub_sweep_parity(ctx, oracle, nc, storage, prec):
    """Every (parity, mu) launch separately, heatbath and overrelaxation."""
    dims = (4, 4, 6, 8)
    beta = BETAS[nc]
    T = thermalised(oracle, nc, dims, storage)
    with make(ctx, nc, dims, beta, storage, prec) as lat:
        for parity in (0, 1):
            for mu in range(4):
                U = load_config(lat, T)
                lat.heatbath_sub(9, 3, parity, mu)
                G = lat.to_host()
                oracle.heatbath_sub(U, dims, beta, 9, 3, parity, mu, storage)
                d = per_link_diff(G, U)
                assert int((d > tol(prec)).sum()) <= (0 if prec == 8 else 2)
                U = load_config(lat, T)
                lat.overrelax_sub(parity, mu)
                G = lat.to_host()
                oracle.overrelax_sub(U, dims, parity, mu, storage)
                assert maxdiff(G, U) <= tol(prec)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_reunitarize_parity(ctx, oracle, nc, storage, prec):
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(3)
    O = oracle.hot_start(nc, dims, 8, 18)
    if storage != 8:   # an 8-real link is unitary by construction; perturb the others
        O = O + 1e-3 * (rng.standard_normal(O.shape) + 1j * rng.standard_normal(O.shape))
    with make(ctx, nc, dims, 6.0, storage, prec) as lat:
        lat.from_host(O)
        U = lat.to_host()
        lat.reunitarize()
        G = lat.to_host()
    oracle.reunitarize(U, storage)
    assert maxdiff(G, U) <= store_tol(storage, prec)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_observables_parity(ctx, oracle, nc, storage, prec):
    dims = (6, 4, 4, 8)
    with make(ctx, nc, dims, 6.0, storage, prec) as lat:
        lat.hot_start(4)
        U = lat.to_host()
        for _ in range(2):
            lat.overrelax()
        U2 = lat.to_host()
        p, pl = lat.plaquette(), lat.polyakov()
    po = oracle.plaquette(U2, dims)
    plo = oracle.polyakov(U2, dims)
    t = 1e-13 if prec == 8 else 2e-6
    assert np.abs(p - po).max() <= t
    assert abs(pl - plo) <= t
