"""sun_wilson_loops_noalloc: the allocation-free form of sun_wilson_loops (line field
and partial sums in the library's fixed device scratch array, results to the
caller's device buffer) gives the allocating entry point's results bit for bit, on
two calls in a row, the second on a smaller lattice (so it runs over what the first
left in the scratch array), and refuses a lattice whose line field does not fit.
The caller's device buffer comes from the HIP runtime directly (the test allocates;
the entry point does not)."""
import ctypes

import numpy as np
import pytest

import sun_amd

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def noalloc_loops(lat, rmax, tmax, stream):
    n = 2 * (rmax + 1) * (tmax + 1)
    host = np.full(n, np.nan)
    p = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(p), host.nbytes) == 0
    try:
        # NaN in the output buffer first, so a slot the call does not write shows
        assert hip().hipMemcpy(p, host.ctypes.data, host.nbytes, 1) == 0      # host to device
        lat.wilson_loops_noalloc(rmax, tmax, p.value, stream)
        sun_amd.sync(stream)
        assert hip().hipMemcpy(host.ctypes.data, p, host.nbytes, 2) == 0      # device to host
    finally:
        hip().hipFree(p)
    return (host[0::2] + 1j * host[1::2]).reshape(rmax + 1, tmax + 1)


@pytest.mark.parametrize("nc,storage,prec", [(3, 18, 8), (3, 12, 4), (2, 18, 8), (4, 18, 4)],
                         ids=["su3_s18_f64", "su3_s12_f32", "su2_f64", "su4_f32"])
def test_noalloc_equals_allocating_entry_point(ctx, nc, storage, prec):
    stream = sun_amd.new_stream()
    try:
        beta = {2: 2.3, 3: 6.0, 4: 11.0}[nc]
        for dims, rmax, tmax, seed in [((8, 8, 8, 12), 4, 6, 3), ((4, 6, 4, 8), 2, 4, 5)]:
            with sun_amd.Lattice(ctx, nc, dims, beta, storage, prec) as lat:
                lat.hot_start(seed)
                lat.generate(seed=seed, n_iter=2, n_or=2, reunit_every=1)
                ref = lat.wilson_loops(rmax, tmax, stream)
                got = noalloc_loops(lat, rmax, tmax, stream)
                assert np.array_equal(got, ref), (dims, np.abs(got - ref).max())
                assert np.abs(got[:, 0] - 1).max() < 1e-5  # W(r, 0) = 1
    finally:
        lib = sun_amd.lib()
        lib.sun_stream_destroy(stream)


def test_noalloc_refuses_what_does_not_fit(ctx):
    """16^3 x 32 SU(3) in FP64: 56.6 MB of Wilson lines, beyond the 32 MiB array"""
    with sun_amd.Lattice(ctx, 3, (16, 16, 16, 32), 6.0, 18, 8) as lat:
        lat.cold_start()
        with pytest.raises(sun_amd.SunError) as e:
            lat.wilson_loops_noalloc(2, 2, 1 << 20)
        assert e.value.code == sun_amd.ERR_BUDGET
        with pytest.raises(sun_amd.SunError) as e:  # misaligned output
            lat.wilson_loops_noalloc(2, 2, (1 << 20) + 8)
        assert e.value.code == sun_amd.ERR_INVALID
