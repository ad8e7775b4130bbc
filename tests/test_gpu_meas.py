"""The measurement chain on the GPU against the FP64 oracle: APE smearing
(ape.cu), stout smearing (stout.cu, SU(3) variants), HYP smearing (hyp.cu; the
device computes each decoration level once, the oracle recurses as the reference
does), multihit (multihit.cu), extended multihit (multihitext.cu) and Wilson loops
(wilsonloop.cu), every variant.

Bars: FP64 1e-12 per link (smearing, multihit) and 1e-13 on W(r, T); FP32 5e-5 per
link plus four times the storage round trip of the oracle's result (8-real FP32
re-derives phases) and, for smearing, ten times the oracle's own change under a
one-ulp input perturbation (iterated projections and the 8-real decode are
ill-conditioned at some links), 2e-6 on W(r, T); HYP in FP32 1e-4 per link (three
nested levels of iterated FP32 projections).  FP32 multihit runs heatbath chains in FP32, so an accept/reject branch may
flip against the oracle; as for heatbath sweeps, every temporal link beyond the bar
must have a near-tie sampler decision among its own updates (flip_report).  The whole chain of
test.cpp:170-230 (APE 25 steps, multihit, Wilson loops) runs in FP64 against the
oracle's chain to 1e-10, and in FP32-12 against it to 2e-3.
"""
import numpy as np
import pytest

import sun_amd
from test_gpu_parity import (BETAS, IDS, MAX_FLIP_OUTLIERS, VARIANTS, flip_report, load_config, per_link_diff,
                             store_tol, thermalised, tol)

pytestmark = pytest.mark.gpu

DIMS = (4, 6, 4, 8)


def smear_bar(lat, U, ref, storage, prec, fn, base=None):
    """Per-link bar for a smearing of input U against the oracle's result ref:
    tol(prec) plus four times the storage round trip of ref; in FP32 also ten times
    the oracle's own change at the link when its input moves by one FP32 ulp
    (relative 2^-23): the device smears in FP32 arithmetic, and the iterated SU(N)
    projections (and the 8-real decode, which divides by |e01|^2+|e02|^2) amplify
    rounding at some links.  fn(P) is the oracle's smearing of P.  Leaves ref on the
    lattice."""
    lat.from_host(ref)
    bar = (tol(prec) if base is None else base) + 4 * per_link_diff(lat.to_host(), ref)
    if prec == 4:
        rng = np.random.default_rng(321)
        P = U * (1 + 2.0 ** -23 * rng.standard_normal(U.shape))
        bar = bar + 10 * per_link_diff(fn(P), ref)
    return bar


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("time_dir", [False, True], ids=["space", "time"])
def test_ape_parity(ctx, oracle, nc, storage, prec, time_dir):
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        lat.ape_smear(0.2, 2, 10, 1e-10, time_dir)
        got = lat.to_host()
        fn = lambda P: oracle.ape_smear(P.copy(), DIMS, 0.2, 2, 10, 1e-10, time_dir)  # noqa: E731
        ref = fn(U)
        bar = smear_bar(lat, U, ref, storage, prec, fn)
    d = per_link_diff(got, ref)
    assert np.all(d <= bar), (float(d.max()), float(bar.max()))
    untouched = slice(0, 3) if time_dir else slice(3, 4)
    assert np.array_equal(got[untouched], U[untouched])


SU3 = [(v, i) for v, i in zip(VARIANTS, IDS) if v[0] == 3]


@pytest.mark.parametrize("nc,storage,prec", [v for v, _ in SU3], ids=[i for _, i in SU3])
@pytest.mark.parametrize("time_dir", [False, True], ids=["space", "time"])
def test_stout_parity(ctx, oracle, nc, storage, prec, time_dir):
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        lat.stout_smear(0.15, 3, time_dir)
        got = lat.to_host()
        fn = lambda P: oracle.stout_smear(P.copy(), DIMS, 0.15, 3, time_dir)  # noqa: E731
        ref = fn(U)
        bar = smear_bar(lat, U, ref, storage, prec, fn)
    d = per_link_diff(got, ref)
    assert np.all(d <= bar), (float(d.max()), float(bar.max()))
    untouched = slice(0, 3) if time_dir else slice(3, 4)
    assert np.array_equal(got[untouched], U[untouched])


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("time_dir", [False, True], ids=["space", "time"])
def test_hyp_parity(ctx, oracle, nc, storage, prec, time_dir):
    """the level-by-level device HYP against the oracle's recursion (hyp.cu)"""
    dims = (4, 4, 4, 6)
    with sun_amd.Lattice(ctx, nc, dims, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, dims, storage))
        lat.hyp_smear((0.75, 0.6, 0.3), 1, 10, 1e-10, time_dir)
        got = lat.to_host()
        fn = lambda P: oracle.hyp_smear(P.copy(), dims, (0.75, 0.6, 0.3), 1, 10, 1e-10, time_dir)  # noqa: E731
        ref = fn(U)
        # FP32: three nested levels of ten-hit SU(N) projections, each rounding in
        # FP32 -- 1e-4 per link (SU(4) links reach 5.3e-5)
        bar = smear_bar(lat, U, ref, storage, prec, fn, base=1e-12 if prec == 8 else 1e-4)
    d = per_link_diff(got, ref)
    assert np.all(d <= bar), (float(d.max()), float(bar.max()))
    untouched = slice(0, 3) if time_dir else slice(3, 4)
    assert np.array_equal(got[untouched], U[untouched])


@pytest.mark.parametrize("nc", [2, 4])
def test_stout_su3_only(ctx, nc):
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], 18, 8) as lat:
        lat.cold_start()
        with pytest.raises(sun_amd.SunError) as e:
            lat.stout_smear(0.1, 1)
        assert e.value.code == sun_amd.ERR_UNSUPPORTED


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_multihit_parity(ctx, oracle, nc, storage, prec):
    beta, seed, nhit = BETAS[nc], 71, 10
    with sun_amd.Lattice(ctx, nc, DIMS, beta, storage, prec) as lat, \
            sun_amd.Lattice(ctx, nc, DIMS, beta, 18, prec) as out:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        lat.multihit(out, seed, nhit)
        got = out.to_host()
    ref, margins = oracle.multihit_margins(U, DIMS, beta, seed, nhit)
    # spatial links copied (decoded in the link precision on both sides)
    assert np.abs(got[:3] - U[:3]).max() <= (0.0 if prec == 8 else store_tol(storage, prec))
    d = per_link_diff(got, ref)
    if prec == 8:
        assert d.max() <= 1e-12
    else:
        # FP32 heatbath hits: a temporal link beyond 5e-5 needs a near-tie decision
        # among its own hits (the staples come from the input, so nothing propagates)
        d[:3 * d.size // 4] = 0.0
        n_out, n_tie, unexplained = flip_report(d, np.full(d.size, 5e-5), margins, DIMS, propagate=False)
        print(f"multihit FP32 su{nc} s{storage}: {n_out} temporal links beyond 5e-5, {n_tie} near-tie")
        assert not unexplained and n_out <= MAX_FLIP_OUTLIERS, (n_out, unexplained[:10])


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_multihit_space_parity(ctx, oracle, nc, storage, prec):
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat, \
            sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], 18, prec) as out:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        lat.multihit_space(out, 4)
        got = out.to_host()
        with pytest.raises(sun_amd.SunError):
            lat.multihit_space(out, 0)
    ref = oracle.multihit_space(U, DIMS, 4)
    assert np.abs(got[3] - U[3]).max() <= (0.0 if prec == 8 else store_tol(storage, prec))
    d = per_link_diff(got[:3], ref[:3])
    assert d.max() <= tol(prec), float(d.max())


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_multihit_ext_parity(ctx, oracle, nc, storage, prec):
    """extended multihit (multihitext.cu): nhit checkerboard heatbath sweeps of each
    site's 3^3 box of temporal links; the device runs 16 lanes per site with the box in
    LDS, the oracle the sites one after another"""
    beta, seed, nhit = BETAS[nc], 73, 3
    with sun_amd.Lattice(ctx, nc, DIMS, beta, storage, prec) as lat, \
            sun_amd.Lattice(ctx, nc, DIMS, beta, 18, prec) as out:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        lat.multihit_ext(out, seed, nhit)
        got = out.to_host()
        with pytest.raises(sun_amd.SunError) as e:
            lat.multihit_ext(out, seed, 0)
        assert e.value.code == sun_amd.ERR_INVALID
    ref, margins = oracle.multihit_ext_margins(U, DIMS, beta, seed, nhit)
    assert np.abs(got[:3] - U[:3]).max() <= (0.0 if prec == 8 else store_tol(storage, prec))
    d = per_link_diff(got, ref)
    if prec == 8:
        assert d.max() <= 1e-12, float(d.max())
    else:
        # FP32 heatbath chains: a temporal link beyond 5e-5 needs a near-tie decision
        # somewhere in its box's updates
        d[:3 * d.size // 4] = 0.0
        n_out, n_tie, unexplained = flip_report(d, np.full(d.size, 5e-5), margins, DIMS, propagate=False)
        print(f"multihit_ext FP32 su{nc} s{storage}: {n_out} temporal links beyond 5e-5, {n_tie} near-tie")
        assert not unexplained and n_out <= MAX_FLIP_OUTLIERS, (n_out, unexplained[:10])


@pytest.mark.parametrize("dims", [(2, 2, 2, 4), (2, 4, 6, 2), (6, 2, 4, 4)])
def test_multihit_ext_small_extents(ctx, oracle, dims):
    """extents of 2 (box offsets -1 and +1 are one physical link, kept as two box
    entries as the reference's padded copy does) and mixed extents, FP64 SU(3)"""
    with sun_amd.Lattice(ctx, 3, dims, 5.8, 18, 8) as lat, sun_amd.Lattice(ctx, 3, dims, 5.8, 18, 8) as out:
        lat.hot_start(9)
        U = lat.to_host()
        lat.multihit_ext(out, 11, 2)
        got = out.to_host()
    ref = oracle.multihit_ext(U, dims, 5.8, 11, 2)
    assert per_link_diff(got, ref).max() <= 1e-12


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_wilson_loops_parity(ctx, oracle, nc, storage, prec):
    rmax, tmax = 3, 5
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        got = lat.wilson_loops(rmax, tmax)
    ref = oracle.wilson_loops(U, DIMS, rmax, tmax)
    assert np.abs(got - ref).max() <= (1e-13 if prec == 8 else 2e-6), np.abs(got - ref).max()
    assert np.abs(got[:, 0] - 1).max() < (1e-13 if prec == 8 else 1e-6)


def test_wilson_loops_odd_volume_and_extent_limits(ctx, oracle):
    dims = (6, 6, 4, 10)  # V = 1440: not a multiple of the 256-site workgroup
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat:
        lat.hot_start(4)
        U = lat.to_host()
        got = lat.wilson_loops(4, 10)
        assert np.abs(got - oracle.wilson_loops(U, dims, 4, 10)).max() <= 1e-13
        for bad in ((5, 2), (2, 11), (-1, 0)):
            with pytest.raises(sun_amd.SunError) as e:
                lat.wilson_loops(*bad)
            assert e.value.code == sun_amd.ERR_INVALID


def test_multihit_refuses_bad_outputs(ctx):
    with sun_amd.Lattice(ctx, 3, DIMS, 6.0, 12, 4) as lat, sun_amd.Lattice(ctx, 3, DIMS, 6.0, 12, 4) as o12, \
            sun_amd.Lattice(ctx, 3, DIMS, 6.0, 18, 8) as o64, sun_amd.Lattice(ctx, 3, (4, 6, 4, 6), 6.0, 18, 4) as od:
        lat.cold_start()
        for out in (lat, o12, o64, od):
            for call in (lambda: lat.multihit(out, 1, 3), lambda: lat.multihit_ext(out, 1, 3)):
                with pytest.raises(sun_amd.SunError) as e:
                    call()
                assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError):
            lat.ape_smear(0.2, -1, 10)


@pytest.mark.parametrize("storage,prec,bar", [(18, 8, 1e-10), (12, 4, 2e-3)], ids=["f64_18", "f32_12"])
def test_static_potential_chain(ctx, oracle, storage, prec, bar):
    """test.cpp:170-230: APE in space (w 0.2, 25 steps, 10 hits), multihit, Wilson
    loops -- GPU chain against the oracle's FP64 chain"""
    dims, beta, seed = (8, 8, 8, 8), 6.0, 5
    rmax, tmax = 4, 4
    U0 = thermalised(oracle, 3, dims, storage)
    with sun_amd.Lattice(ctx, 3, dims, beta, storage, prec) as lat, \
            sun_amd.Lattice(ctx, 3, dims, beta, 18, prec) as mh:
        U = load_config(lat, U0)
        lat.ape_smear(0.2, 25, 10, 1e-10)
        lat.multihit(mh, seed, 20)
        got = mh.wilson_loops(rmax, tmax)
    ref_s = oracle.ape_smear(U.copy(), dims, 0.2, 25, 10, 1e-10)
    ref = oracle.wilson_loops(oracle.multihit(ref_s, dims, beta, seed, 20), dims, rmax, tmax)
    assert np.abs(got - ref).max() <= bar, np.abs(got - ref).max()
    # smearing and multihit make the large loops measurable: W(4, 4) well above noise
    assert got[4, 4].real > 1e-3


@pytest.mark.parametrize("ss", range(6), ids=["plain", "ape_s", "mhit", "ape_t", "ape_ts", "mhit_ape_s"])
def test_reference_measurement_variants(ctx, oracle, ss):
    """test.cpp:170-215: the six variants the reference's test program measures on every
    configuration -- none, APE in space (w 0.2, 25 steps, 10 hits), multihit (100 hits),
    APE in time (0.2, 1 step), APE in time then in space, multihit then APE in space
    (0.4, 50 steps) -- each followed by W(r, T), r <= X/2, T <= T/2
    (test.cpp:217-221); FP64 SU(3) 8^4 against the oracle's chain"""
    dims, beta, seed = (8, 8, 8, 8), 6.0, 6
    rmax, tmax = 4, 4
    U0 = thermalised(oracle, 3, dims, 18)
    with sun_amd.Lattice(ctx, 3, dims, beta, 18, 8) as lat, sun_amd.Lattice(ctx, 3, dims, beta, 18, 8) as mh:
        U = load_config(lat, U0)
        ref = U.copy()
        meas = lat
        if ss == 1:
            lat.ape_smear(0.2, 25, 10, 1e-10)
            oracle.ape_smear(ref, dims, 0.2, 25, 10, 1e-10)
        elif ss in (2, 5):
            lat.multihit(mh, seed, 100)
            ref = oracle.multihit(ref, dims, beta, seed, 100)
            meas = mh
            if ss == 5:
                mh.ape_smear(0.4, 50, 10, 1e-10)
                oracle.ape_smear(ref, dims, 0.4, 50, 10, 1e-10)
        elif ss in (3, 4):
            lat.ape_smear(0.2, 1, 10, 1e-10, time_dir=True)
            oracle.ape_smear(ref, dims, 0.2, 1, 10, 1e-10, time_dir=True)
            if ss == 4:
                lat.ape_smear(0.2, 25, 10, 1e-10)
                oracle.ape_smear(ref, dims, 0.2, 25, 10, 1e-10)
        got = meas.wilson_loops(rmax, tmax)
    want = oracle.wilson_loops(ref, dims, rmax, tmax)
    assert np.abs(got - want).max() <= 1e-10, np.abs(got - want).max()
    assert np.abs(got[:, 0] - 1).max() < 1e-12  # W(r, 0) = 1: the spatial links stay in SU(3)
