"""Link monitors on the GPU against the oracle (linktrsum.cu:130-150,
linkdetsum.cu:84-101).

The device reads the stored links (rows reconstructed as the update kernels read
them) and does FP64 arithmetic; the oracle gets the same decoded field from
sun_links_to_host.  Mean trace and determinant agree to 1e-13 (sums of O(1) values
in another order), the maximal deviations to 1e-14 absolute, on thermalised and
Haar fields of every variant.  The identity field gives (1, 0, 1, 0, 0, 0) exactly;
GaugeTrace/GaugeDet entry points and the asynchronous form return the same numbers.
A drift test runs FP32 chains without reunitarisation and checks that the monitor
sees the drift and that reunitarisation removes it.
"""
import numpy as np
import pytest

import sun_amd
from test_gpu_parity import BETAS, IDS, VARIANTS, load_config, thermalised

pytestmark = pytest.mark.gpu


def check(got, ref):
    assert np.abs(got[:4] - ref[:4]).max() <= 1e-13, (got, ref)
    assert np.abs(got[4:] - ref[4:]).max() <= 1e-14, (got, ref)


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_monitor_parity(ctx, oracle, nc, storage, prec):
    dims = (4, 6, 4, 8)
    with sun_amd.Lattice(ctx, nc, dims, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, thermalised(oracle, nc, dims, storage))
        check(lat.link_monitor(), oracle.link_monitor(U))
        lat.hot_start(3)
        H = lat.to_host()
        got = lat.link_monitor()
        check(got, oracle.link_monitor(H))
        assert lat.link_trace() == complex(got[0], got[1])
        assert lat.link_det() == complex(got[2], got[3])


@pytest.mark.parametrize("nc,storage,prec", [v for v in VARIANTS if v[1] != 8],
                         ids=[i for v, i in zip(VARIANTS, IDS) if v[1] != 8])
def test_monitor_identity_exact(ctx, nc, storage, prec):
    with sun_amd.Lattice(ctx, nc, (4, 4, 4, 4), 6.0, storage, prec) as lat:
        lat.cold_start()
        assert list(lat.link_monitor()) == [1.0, 0.0, 1.0, 0.0, 0.0, 0.0]


def test_monitor_odd_volume_and_async(ctx, oracle):
    # V = 6*6*4*10 = 1440: not a multiple of the 256-site workgroup
    torch = pytest.importorskip("torch")
    dims = (6, 6, 4, 10)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as lat:
        lat.hot_start(9)
        ref = oracle.link_monitor(lat.to_host())
        got = lat.link_monitor()
        check(got, ref)
        dev = torch.full((8,), -7.0, dtype=torch.float64, device="cuda")
        rc = sun_amd.lib().sun_link_monitor_async(lat.h, None, dev.data_ptr())
        assert rc == 0
        pl = torch.zeros(2, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()  # the fills run on torch's stream
        assert sun_amd.lib().sun_plaquette_async(lat.h, None, pl.data_ptr()) == 0
        lat.link_monitor()  # synchronises the lattice's stream
        torch.cuda.synchronize()
        out = dev.cpu().numpy()
        assert np.array_equal(out[:6], got) and list(out[6:]) == [-7.0, -7.0]
        assert np.array_equal(pl.cpu().numpy(), lat.plaquette())


@pytest.mark.parametrize("storage", [18, 12])
def test_fp32_drift_seen_and_removed(ctx, storage):
    dims = (8, 8, 8, 8)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, storage, 4) as lat:
        lat.hot_start(5)
        lat.reunitarize()
        start = lat.link_monitor()
        for sweep in range(20):
            lat.heatbath(seed=17, sweep=sweep)
            for _ in range(4):
                lat.overrelax()
        drift = lat.link_monitor()
        lat.reunitarize()
        after = lat.link_monitor()
        assert start[4] < 1e-6 and after[4] < 1e-6
        assert drift[4] > after[4], (drift[4], after[4])
        assert abs(after[2] - 1) < 1e-6 and after[5] < 2e-6


def test_fp64_stays_unitary(ctx):
    with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, 18, 8) as lat:
        lat.hot_start(5)
        for sweep in range(3):
            lat.heatbath(seed=17, sweep=sweep)
            lat.overrelax()
        m = lat.link_monitor()
        assert m[4] < 1e-13 and m[5] < 1e-13
