"""Operator-basis Wilson loops on the GPU against the FP64 oracle: the operator
fields of src/wl/calcop_dg_A0.cu (10 slots) and src/wl/calcop_dg_33.cu (25 slots),
and the correlator of wilsonloop_dg_A0.cu over every slot pair, every variant.

Bars: FP64 operator fields equal the oracle's bit for bit (same links, same product
and sum order); FP64 correlators within 1e-12 (the per-workgroup partial sums are
added in another order than the oracle's chunks).  FP32: 5e-5 per operator entry
(products of up to 2 x 8 + r links in FP32) and 2e-5 on the normalised correlators.
At 32^4 (the reference's Sigma_g_plus(12, 16) measurement, test.cpp:255-298) the
slot-0 pair equals W(r, T) of the plain Wilson loops, the T = 0 matrix is symmetric,
and the correlators are bounded by the operator norms.
"""
import numpy as np
import pytest

import sun_amd
from test_gpu_parity import BETAS, IDS, VARIANTS, load_config, thermalised

pytestmark = pytest.mark.gpu

DIMS = (4, 6, 4, 8)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_wlop_fields_parity(ctx, oracle, nc, storage, prec):
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        for family in (sun_amd.WLOP_A0, sun_amd.WLOP_33):
            with sun_amd.WilsonLoopOps(lat, family, 2) as ops:
                for r, mu in ((2, 0), (3, 2), (1, 1)):
                    ops.fields(r, mu)
                    got = ops.fields_to_host()
                    want = oracle.wlop_fields(U, DIMS, family, r, mu)
                    if prec == 8:
                        assert np.array_equal(got, want), (family, r, mu, float(np.abs(got - want).max()))
                    else:
                        assert np.abs(got - want).max() <= 5e-5, (family, r, mu, float(np.abs(got - want).max()))


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("family", [0, 1], ids=["A0", "33"])
def test_wilson_loops_ops_parity(ctx, oracle, nc, storage, prec, family):
    rmax, tmax = 3, 4
    with sun_amd.Lattice(ctx, nc, DIMS, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, DIMS, storage))
        got = lat.wilson_loops_ops(rmax, tmax, family)
    want = oracle.wilson_loops_ops(U, DIMS, rmax, tmax, family)
    bar = 1e-12 if prec == 8 else 2e-5
    assert np.abs(got - want).max() <= bar, float(np.abs(got - want).max())
    assert np.abs(want).max() > 0.1


def test_wlop_object_api(ctx, oracle):
    dims = (4, 4, 4, 6)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat:
        lat.hot_start(5)
        U = lat.to_host()
        with sun_amd.WilsonLoopOps(lat, sun_amd.WLOP_33, 3) as ops:
            assert ops.nop == 25
            with pytest.raises(sun_amd.SunError) as e:
                ops.correlate(2, 0)
            assert e.value.code == sun_amd.ERR_STATE
            ops.fields(2, 0)
            ops.correlate(2, 0)
            ops.fields(2, 1)
            ops.correlate(2, 1)
            acc = ops.read()
            want = np.zeros_like(acc)
            for mu in (0, 1):
                oracle.wlop_correlate(U, dims, oracle.wlop_fields(U, dims, 1, 2, mu), 2, mu, 3, want)
            assert np.abs(acc - want).max() <= 1e-12 * np.abs(want).max()
            # component blocks (wilsonloop_dg.cu:101-122 order)
            comps, pos = [3, 2], [1, 9]
            blk = ops.read_blocks(comps, pos)
            rows = [(p + a) * 25 + (p + b) for c, p in zip(comps, pos) for a in range(c) for b in range(c)]
            assert np.array_equal(blk, acc[rows])
            with pytest.raises(sun_amd.SunError):
                ops.read_blocks([3], [23])
            ops.clear()
            assert not ops.read().any()


def test_wlop_pair_zero_is_the_wilson_loop(ctx):
    dims = (6, 6, 6, 8)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as lat:
        lat.hot_start(3)
        for it in range(3):
            lat.heatbath(4, it)
        W = lat.wilson_loops(4, 5).real
        for family in (0, 1):
            got = lat.wilson_loops_ops(4, 5, family)
            assert np.abs(got[:, 0, :] - W[2:]).max() < 2e-6, family


def test_wlop_odd_volume_and_limits(ctx, oracle):
    dims = (6, 6, 4, 10)  # V = 1440: not a multiple of the 256-site workgroup
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat:
        lat.hot_start(4)
        U = lat.to_host()
        with sun_amd.WilsonLoopOps(lat, sun_amd.WLOP_A0, 10) as ops:
            for r, mu in ((6, 0), (4, 2)):
                ops.fields(r, mu)
                assert np.array_equal(ops.fields_to_host(), oracle.wlop_fields(U, dims, 0, r, mu))
                ops.clear()
                ops.correlate(r, mu)
                want = oracle.wlop_correlate(U, dims, oracle.wlop_fields(U, dims, 0, r, mu), r, mu, 10)
                assert np.abs(ops.read() - want).max() <= 1e-12 * np.abs(want).max()
            for bad in ((0, 0), (7, 0), (5, 2), (2, 3), (2, -1)):
                with pytest.raises(sun_amd.SunError) as e:
                    ops.fields(*bad)
                assert e.value.code == sun_amd.ERR_INVALID
        for bad in ((2, 11), (2, -1)):
            with pytest.raises(sun_amd.SunError):
                sun_amd.WilsonLoopOps(lat, 0, bad[1])
        with pytest.raises(sun_amd.SunError):
            sun_amd.WilsonLoopOps(lat, 2, 3)
        with pytest.raises(sun_amd.SunError):
            lat.wilson_loops_ops(1, 3)
        with pytest.raises(sun_amd.SunError):
            lat.wilson_loops_ops(5, 3)  # beyond the z extent


def test_wlop_fullsize_32x4(ctx):
    """the reference's measurement (test.cpp:255-298: Sigma_g_plus(12, 16), r = 2..12)
    on a 32^4 SU(3) FP64 field"""
    dims = (32, 32, 32, 32)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat:
        lat.hot_start(11)
        for it in range(2):
            lat.heatbath(12, it)
            lat.overrelax()
        print("32^4 operator-basis Wilson loops: A0, r = 2..12, T <= 16", flush=True)
        W = lat.wilson_loops(12, 16).real
        got = lat.wilson_loops_ops(12, 16, sun_amd.WLOP_A0)
    assert got.shape == (11, 100, 17)
    assert np.abs(got[:, 0, :] - W[2:]).max() < 1e-12
    G0 = got[:, :, 0].reshape(11, 10, 10)
    assert np.abs(G0 - G0.swapaxes(1, 2)).max() < 1e-12
    # |Re Tr(A U B^+ U')| <= |A| |B| (Frobenius, U, U' unitary): the MO0 slots are four
    # SU(3) paths / 2, |O| <= 2 sqrt 3, so a site and direction adds at most 12 and
    # |W_ij| <= 3 * 12 / 9 after the 3 N V norm
    assert np.abs(got).max() <= 4.0 + 1e-12
    assert not G0[:, 9, :].any() and not G0[:, :, 9].any()


@pytest.mark.parametrize("storage,prec", [(18, 8), (12, 4)], ids=["f64-18", "f32-12"])
def test_wlop_33_fullsize_32x4(ctx, storage, prec):
    """the 33 family at the reference's lattice size (32^4 SU(3), r = 12, T <= 16), FP64
    and FP32 (the LDS-staged correlator of both): slot 0 is the line of the A0 family,
    so the (0, 0) correlator equals A0's; the T = 0 matrix is symmetric; and
    |W_ij(T)| <= sqrt(W_ii(0) W_jj(0)) (Cauchy-Schwarz over sites and entries: the
    temporal transport keeps the Frobenius norm of O_i).  FP32: products of up to
    2 x 8 + 12 FP32 links per entry, sums in FP64 (relative bars 1e-5)."""
    dims = (32, 32, 32, 32)
    rel = 1e-12 if prec == 8 else 1e-5
    with sun_amd.Lattice(ctx, 3, dims, 6.0, storage, prec) as lat:
        lat.hot_start(13)
        for it in range(2):
            lat.heatbath(14, it)
            lat.overrelax()
        print("32^4 operator-basis correlators: 33 and A0 families, r = 12, mu = 1, T <= 16", flush=True)
        got = {}
        for family in (sun_amd.WLOP_33, sun_amd.WLOP_A0):
            with sun_amd.WilsonLoopOps(lat, family, 16) as ops:
                ops.fields(12, 1)
                ops.correlate(12, 1)
                got[family] = ops.read().reshape(ops.nop, ops.nop, 17)
    G = got[sun_amd.WLOP_33]
    assert np.abs(G[0, 0] - got[sun_amd.WLOP_A0][0, 0]).max() <= rel * np.abs(G[0, 0]).max()
    G0 = G[:, :, 0]
    assert np.abs(G0 - G0.T).max() <= rel * np.abs(G0).max()
    d = np.sqrt(np.diag(G0))
    assert (np.diag(G0) > 0).all()
    assert (np.abs(G) <= np.multiply.outer(d, d)[:, :, None] * (1 + rel)).all()
    assert np.abs(G[:, :, 16]).max() > 0
