"""Edge cases on the GPU against the oracle: the smallest lattices (2^4, where every
row of a parity is one site and no brick shape applies, and 2x4x2x6), for sweeps,
observables, reunitarisation and the measurement chain, in FP64 (1e-12) for every N;
refusal of invalid shapes at creation; and a lattice whose final sums take whole
passes of batched loads and a tail (32^3 x 40)."""
import numpy as np
import pytest

import sun_amd
from test_gpu_parity import BETAS, load_config, thermalised

pytestmark = pytest.mark.gpu

TINY = [(2, 2, 2, 2), (2, 4, 2, 6)]


@pytest.mark.parametrize("dims", TINY)
@pytest.mark.parametrize("nc", [2, 3, 4])
def test_tiny_lattice_sweeps_and_observables(ctx, oracle, nc, dims):
    beta = BETAS[nc]
    with sun_amd.Lattice(ctx, nc, dims, beta, 18, 8) as lat:
        U = load_config(lat, thermalised(oracle, nc, dims, 18))
        ref = U.copy()
        for sweep in range(2):
            lat.heatbath(seed=3, sweep=sweep)
            oracle.heatbath_sweep(ref, dims, beta, 3, sweep)
            lat.overrelax()
            oracle.overrelax_sweep(ref, dims)
        assert np.abs(lat.to_host() - ref).max() < 1e-12
        lat.reunitarize()
        oracle.reunitarize(ref)
        G = lat.to_host()
        assert np.abs(G - ref).max() < 1e-12
        assert np.abs(lat.plaquette() - oracle.plaquette(G, dims)).max() < 1e-13
        assert abs(lat.polyakov() - oracle.polyakov(G, dims)) < 1e-13
        W = lat.wilson_loops(1, dims[3])
        assert np.abs(W - oracle.wilson_loops(G, dims, 1, dims[3])).max() < 1e-13
        lat.ape_smear(0.2, 1, 10, 1e-10)
        assert np.abs(lat.to_host() - oracle.ape_smear(G.copy(), dims, 0.2, 1, 10, 1e-10)).max() < 1e-12


@pytest.mark.parametrize("dims", [(3, 4, 4, 4), (4, 4, 4, 0), (4, 4, 4, 1)])
def test_invalid_extents_refused(ctx, dims):
    with pytest.raises(sun_amd.SunError) as e:
        sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8)
    assert e.value.code == sun_amd.ERR_INVALID


def test_final_sums_whole_passes_and_tail(ctx, oracle):
    """32^3 x 40: 5120 plaquette and link-monitor partials, so each thread of the
    fixed-order final sums takes two whole passes of eight loads and a tail, and every
    time line of the Polyakov loop has 10 links per segment.  FP64-18 Haar field
    against the oracle (a partial dropped or summed twice moves the plaquette by
    ~1e-6 here); then one generation iteration, whose fused measurement (one grid and
    one finalize workgroup) must equal the separate plaquette and Polyakov kernels bit
    for bit and the oracle on the field it leaves."""
    dims = (32, 32, 32, 40)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat:
        lat.hot_start(11)
        H = lat.to_host()
        assert np.abs(lat.plaquette() - oracle.plaquette(H, dims)).max() <= 1e-13
        assert abs(lat.polyakov() - oracle.polyakov(H, dims)) <= 1e-13
        mon, ref = lat.link_monitor(), oracle.link_monitor(H)
        assert np.abs(mon[:4] - ref[:4]).max() <= 1e-13 and np.abs(mon[4:] - ref[4:]).max() <= 1e-14
        del H
        obs = lat.generate(seed=3, n_iter=1)
        p, pl = lat.plaquette(), lat.polyakov()
        assert list(obs[0]) == [p[0], p[1], pl.real, pl.imag]
        G = lat.to_host()
        assert np.abs(p - oracle.plaquette(G, dims)).max() <= 1e-13
        assert abs(pl - oracle.polyakov(G, dims)) <= 1e-13
