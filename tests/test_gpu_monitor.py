"""Link monitors on the GPU against the oracle (linktrsum.cu:130-150,
linkdetsum.cu:84-101).

The device reads the stored links (rows reconstructed as the update kernels read
them) and does FP64 arithmetic; the oracle gets the same decoded field from
sun_links_to_host.  Mean trace and determinant agree to 1e-13 (sums of O(1) values
in another order), the maximal deviations to 1e-14 absolute, on thermalised and
Haar fields of every variant.  The identity field gives (1, 0, 1, 0, 0, 0) exactly;
GaugeTrace/GaugeDet entry points and the asynchronous form return the same numbers.
Drift tests run FP32 chains without reunitarisation: the monitor sees the drift,
reunitarisation removes it, and the measured growth (geometric for 12-real storage,
slow for 18-real) is what sets the reunitarisation interval.  A NaN anywhere in a
link makes the determinant and deviation outputs NaN (never a silent 0).
"""
import ctypes

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


class DeviceBuffer:
    """a few bytes of device memory through the HIP runtime (test plumbing)"""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.p, self.n = ctypes.c_void_p(), nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def get(self, n):
        a = np.zeros(n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_monitor_odd_volume_and_async(ctx, oracle):
    # V = 6*6*4*10 = 1440: not a multiple of the 256-site workgroup
    dims = (6, 6, 4, 10)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as lat:
        lat.hot_start(9)
        ref = oracle.link_monitor(lat.to_host())
        got = lat.link_monitor()
        check(got, ref)
        buf = DeviceBuffer(80)
        try:
            buf.put(np.full(10, -7.0))
            assert sun_amd.lib().sun_link_monitor_async(lat.h, None, buf.p) == 0
            lat.link_monitor()  # same stream: the asynchronous result is written by now
            out = buf.get(10)
            assert np.array_equal(out[:6], got) and list(out[6:]) == [-7.0] * 4
            assert sun_amd.lib().sun_plaquette_async(lat.h, None, buf.p) == 0
            assert sun_amd.lib().sun_polyakov_async(lat.h, None, ctypes.c_void_p(buf.p.value + 16)) == 0
            lat.link_monitor()
            out = buf.get(10)
            p = lat.plaquette()
            q = lat.polyakov()
            assert list(out[:4]) == [p[0], p[1], q.real, q.imag]
        finally:
            buf.free()


def test_monitor_reports_nan(ctx, oracle):
    with sun_amd.Lattice(ctx, 3, (4, 4, 4, 4), 6.0, 18, 8) as lat:
        U = oracle.cold_start(3, (4, 4, 4, 4))
        U[1, 77, 0, 1] = np.nan
        lat.from_host(U)
        got = lat.link_monitor()
        # an off-diagonal entry: the trace stays finite, everything else is NaN
        assert np.isfinite(got[:2]).all() and np.isnan(got[2:]).all()


def iteration(lat, sweep):
    lat.heatbath(seed=17, sweep=sweep)
    for _ in range(4):
        lat.overrelax()


@pytest.mark.parametrize("storage", [18, 12])
def test_fp32_drift_seen_and_removed(ctx, storage):
    with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, storage, 4) as lat:
        lat.hot_start(5)
        lat.reunitarize()
        start = lat.link_monitor()
        for sweep in range(3):
            iteration(lat, sweep)
        drift = lat.link_monitor()
        lat.reunitarize()
        after = lat.link_monitor()
        assert start[4] < 1e-6 and after[4] < 1e-6 and after[5] < 1e-6
        assert drift[4] > 10 * after[4], (drift[4], after[4])


def test_fp32_reunitarisation_interval(ctx):
    """Without reunitarisation the 12-real FP32 deviation grows geometrically (about
    4x per iteration of 1 HB + 4 OR sweeps: row 2 is rebuilt from rows 0 and 1, so
    their error feeds back through every SU(2) update) and the field overflows after
    about a dozen iterations; 18-real FP32 drifts slowly (~1e-5 after 20
    iterations).  The generation driver's default, reunitarisation every iteration,
    keeps both at the FP32 rounding level."""
    dev = {}
    for storage in (12, 18):
        with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, storage, 4) as lat:
            lat.hot_start(5)
            lat.reunitarize()
            dev[storage] = []
            for sweep in range(6):
                iteration(lat, sweep)
                dev[storage].append(lat.link_monitor()[4])
    d12, d18 = np.array(dev[12]), np.array(dev[18])
    assert np.all(d12[1:] / d12[:-1] > 2.0), d12
    assert d12[-1] > 1e-3 and d18[-1] < 2e-5, (d12, d18)
    with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, 12, 4) as lat:
        lat.hot_start(5)
        lat.generate(seed=17, n_iter=6, n_or=4, reunit_every=1)
        assert lat.link_monitor()[4] < 1e-6


def test_fp64_stays_unitary(ctx):
    with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, 18, 8) as lat:
        lat.hot_start(5)
        for sweep in range(3):
            lat.heatbath(seed=17, sweep=sweep)
            lat.overrelax()
        m = lat.link_monitor()
        assert m[4] < 1e-13 and m[5] < 1e-13
