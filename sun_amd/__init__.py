"""sun_amd -- Python binding of the MI355X SU(N) quenched ensemble-generation engine.

Everything here calls the C ABI of ``libsun_hip.so`` (declared in ``include/sun.h``)
through ctypes; there is no fallback.  If the library is missing or does not load,
``lib()`` raises: build it with ``python sun_amd/build.py`` (or
``__graft_entry__.build()``).

The classes mirror the reference's host API (namespace CULQCD):

=====================  =========================================================
reference              here
=====================  =========================================================
initCULQCD/EndCULQCD   ``Context`` (``sun_init`` / ``sun_finalize``)
SETPARAMS + _gauge     ``Lattice`` (``sun_lattice_create`` / ``destroy``)
gauge.Init()           ``Lattice.cold_start()``
gauge.Init(RNG)        ``Lattice.hot_start(seed)``
HeatBath(...).Run()    ``Lattice.heatbath(seed, sweep)``
OverRelaxation.Run()   ``Lattice.overrelax()``
Reunitarize.Run()      ``Lattice.reunitarize()``
PlaquetteCUB.Run()     ``Lattice.plaquette()``  -> (spatial, temporal)
OnePolyakovLoop.Run()  ``Lattice.polyakov()``   -> complex
=====================  =========================================================
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsun_hip.so")

STORE_18, STORE_12, STORE_8 = 18, 12, 8
FILE_NATIVE, FILE_REFERENCE, FILE_REFERENCE_NOHEADER = 0, 1, 2
F32, F64 = 4, 8
_lib = None

EXPORTS = [
    "sun_init", "sun_finalize", "sun_last_error", "sun_version",
    "sun_lattice_create", "sun_lattice_destroy", "sun_lattice_bytes", "sun_lattice_copy",
    "sun_lattice_convert",
    "sun_lattice_get_launch", "sun_lattice_set_launch", "sun_ctx_set_launch", "sun_lattice_save", "sun_lattice_load",
    "sun_cold_start", "sun_hot_start",
    "sun_heatbath", "sun_overrelax", "sun_heatbath_sub", "sun_overrelax_sub", "sun_reunitarize",
    "sun_plaquette", "sun_polyakov", "sun_plaquette_async", "sun_polyakov_async",
    "sun_link_monitor", "sun_link_monitor_async", "sun_link_trace", "sun_link_det",
    "sun_ape_smear", "sun_stout_smear", "sun_hyp_smear", "sun_multihit", "sun_multihit_space", "sun_multihit_ext",
    "sun_wilson_loops", "sun_wilson_loops_noalloc",
    "sun_wlop_create", "sun_wlop_destroy", "sun_wlop_nops", "sun_wlop_fields", "sun_wlop_correlate",
    "sun_wlop_clear", "sun_wlop_read", "sun_wlop_read_blocks", "sun_wlop_fields_to_host", "sun_wilson_loops_ops",
    "sun_links_to_host", "sun_links_from_host",
    "sun_generate",
    "sun_batch_create", "sun_batch_destroy", "sun_batch_get_launch", "sun_batch_set_launch",
    "sun_batch_heatbath", "sun_batch_overrelax",
    "sun_batch_reunitarize", "sun_batch_plaquette", "sun_batch_generate",
    "sun_server_create", "sun_server_destroy", "sun_server_submit", "sun_server_cancel",
    "sun_server_step", "sun_server_start", "sun_server_stop", "sun_server_wait", "sun_server_status",
    "sun_server_observables", "sun_server_release", "sun_server_stats", "sun_server_clock",
    "sun_server_suspend", "sun_server_resume", "sun_server_job_error", "sun_job_bytes",
    "sun_stream_create", "sun_stream_destroy", "sun_stream_synchronize",
    "sun_event_create", "sun_event_destroy", "sun_event_record", "sun_event_elapsed_ms",
    "sun_device_name",
]


ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_BUDGET, ERR_STATE = -1, -2, -3, -4, -5, -6


class SunError(RuntimeError):
    """A failed call; `code` is the SUN_ERR_* status it returned."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class LatticeDesc(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("dims", ctypes.c_int * 4), ("beta", ctypes.c_double),
                ("storage", ctypes.c_int), ("precision", ctypes.c_int)]


class GenerateDesc(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("first_sweep", ctypes.c_uint32), ("n_iter", ctypes.c_int),
                ("n_or", ctypes.c_int), ("reunit_every", ctypes.c_int), ("measure", ctypes.c_int),
                ("use_graph", ctypes.c_int)]


class Launch(ctypes.Structure):
    """sun_launch: block, brick (log2 rows along x1, x2, x3; brick[0] < 0 = linear),
    lds (unused dynamic LDS per workgroup), piped (0 plain, 1/2 software-pipelined by
    directions / half-directions)"""
    _fields_ = [("block", ctypes.c_int), ("brick", ctypes.c_int * 3), ("lds", ctypes.c_int),
                ("piped", ctypes.c_int)]

    def as_dict(self):
        return {"block": self.block, "brick": list(self.brick), "lds": self.lds, "piped": self.piped}


class JobDesc(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int), ("dims", ctypes.c_int * 4), ("beta", ctypes.c_double),
                ("storage", ctypes.c_int), ("precision", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("hot_start", ctypes.c_int), ("n_iter", ctypes.c_int), ("n_or", ctypes.c_int),
                ("reunit_every", ctypes.c_int)]


class JobStatus(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("state", ctypes.c_int), ("iters_done", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("plaq", ctypes.c_double * 2), ("poly", ctypes.c_double * 2)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("jobs_done", ctypes.c_uint64), ("link_updates", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("busy_s", ctypes.c_double), ("budget", ctypes.c_size_t), ("bytes_in_use", ctypes.c_size_t),
                ("running", ctypes.c_int), ("pending", ctypes.c_int)]


JOB_PENDING, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_SUSPENDED, JOB_FAILED = 0, 1, 2, 3, 4, 5


def lib() -> ctypes.CDLL:
    """Load libsun_hip.so; raises if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SunError(f"{LIB_PATH} is missing: build it with `python sun_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    u32, u64 = ctypes.c_uint32, ctypes.c_uint64
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "sun_init": (ci, [ci, ci, ci, ctypes.POINTER(vp)]),
        "sun_finalize": (ci, [vp]),
        "sun_last_error": (ctypes.c_char_p, []),
        "sun_version": (ctypes.c_char_p, []),
        "sun_lattice_create": (ci, [vp, ctypes.POINTER(LatticeDesc), ctypes.POINTER(vp)]),
        "sun_lattice_destroy": (ci, [vp]),
        "sun_lattice_bytes": (ctypes.c_size_t, [ctypes.POINTER(LatticeDesc)]),
        "sun_lattice_copy": (ci, [vp, vp, vp]),
        "sun_lattice_convert": (ci, [vp, vp, vp]),
        "sun_ctx_set_launch": (ci, [vp, ctypes.POINTER(LatticeDesc), ci, ctypes.POINTER(Launch)]),
        "sun_lattice_get_launch": (ci, [vp, ci, ctypes.POINTER(Launch)]),
        "sun_lattice_set_launch": (ci, [vp, ci, ctypes.POINTER(Launch)]),
        "sun_lattice_save": (ci, [vp, ctypes.c_char_p, ci, ci, u64, u32]),
        "sun_lattice_load": (ci, [vp, ctypes.c_char_p, ci, ci, ctypes.POINTER(u64), ctypes.POINTER(u32)]),
        "sun_cold_start": (ci, [vp, vp]),
        "sun_hot_start": (ci, [vp, u64, vp]),
        "sun_heatbath": (ci, [vp, u64, u32, vp]),
        "sun_overrelax": (ci, [vp, vp]),
        "sun_heatbath_sub": (ci, [vp, u64, u32, ci, ci, vp]),
        "sun_overrelax_sub": (ci, [vp, ci, ci, vp]),
        "sun_reunitarize": (ci, [vp, vp]),
        "sun_plaquette": (ci, [vp, vp, dp]),
        "sun_polyakov": (ci, [vp, vp, dp]),
        "sun_plaquette_async": (ci, [vp, vp, vp]),
        "sun_polyakov_async": (ci, [vp, vp, vp]),
        "sun_link_monitor": (ci, [vp, vp, dp]),
        "sun_link_monitor_async": (ci, [vp, vp, vp]),
        "sun_link_trace": (ci, [vp, vp, dp]),
        "sun_link_det": (ci, [vp, vp, dp]),
        "sun_ape_smear": (ci, [vp, cd, ci, ci, cd, ci, vp]),
        "sun_stout_smear": (ci, [vp, cd, ci, ci, vp]),
        "sun_hyp_smear": (ci, [vp, cd, cd, cd, ci, ci, cd, ci, vp]),
        "sun_multihit": (ci, [vp, vp, u64, ci, vp]),
        "sun_multihit_space": (ci, [vp, vp, ci, vp]),
        "sun_multihit_ext": (ci, [vp, vp, u64, ci, vp]),
        "sun_wilson_loops": (ci, [vp, ci, ci, dp, vp]),
        "sun_wilson_loops_noalloc": (ci, [vp, ci, ci, vp, vp]),
        "sun_wlop_create": (ci, [vp, ci, ci, ctypes.POINTER(vp)]),
        "sun_wlop_destroy": (ci, [vp]),
        "sun_wlop_nops": (ci, [vp]),
        "sun_wlop_fields": (ci, [vp, ci, ci, vp]),
        "sun_wlop_correlate": (ci, [vp, ci, ci, vp]),
        "sun_wlop_clear": (ci, [vp, vp]),
        "sun_wlop_read": (ci, [vp, dp, vp]),
        "sun_wlop_read_blocks": (ci, [vp, ci, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp, vp]),
        "sun_wlop_fields_to_host": (ci, [vp, dp]),
        "sun_wilson_loops_ops": (ci, [vp, ci, ci, ci, dp, vp]),
        "sun_links_to_host": (ci, [vp, dp, ci]),
        "sun_links_from_host": (ci, [vp, dp, ci]),
        "sun_generate": (ci, [vp, ctypes.POINTER(GenerateDesc), vp, dp]),
        "sun_batch_create": (ci, [vp, ci, ctypes.POINTER(vp), ctypes.POINTER(u64), ctypes.POINTER(vp)]),
        "sun_batch_destroy": (ci, [vp]),
        "sun_batch_get_launch": (ci, [vp, ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]),
        "sun_batch_set_launch": (ci, [vp, ci, ci, ci, ci]),
        "sun_batch_heatbath": (ci, [vp, ctypes.POINTER(u32), vp]),
        "sun_batch_overrelax": (ci, [vp, vp]),
        "sun_batch_reunitarize": (ci, [vp, vp]),
        "sun_batch_plaquette": (ci, [vp, vp, dp]),
        "sun_batch_generate": (ci, [vp, u32, ci, ci, ci, ci, vp, dp]),
        "sun_server_create": (ci, [vp, ctypes.c_size_t, ci, ctypes.POINTER(vp)]),
        "sun_server_destroy": (ci, [vp]),
        "sun_server_submit": (ci, [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(u64)]),
        "sun_server_cancel": (ci, [vp, u64]),
        "sun_server_step": (ci, [vp]),
        "sun_server_start": (ci, [vp]),
        "sun_server_stop": (ci, [vp]),
        "sun_server_wait": (ci, [vp, u64, cd]),
        "sun_server_status": (ci, [vp, u64, ctypes.POINTER(JobStatus)]),
        "sun_server_observables": (ci, [vp, u64, dp, ci]),
        "sun_server_release": (ci, [vp, u64]),
        "sun_server_stats": (ci, [vp, ctypes.POINTER(ServerCounters)]),
        "sun_server_clock": (cd, [vp]),
        "sun_server_suspend": (ci, [vp, u64, ctypes.c_char_p]),
        "sun_server_resume": (ci, [vp, ctypes.c_char_p, ctypes.POINTER(u64)]),
        "sun_server_job_error": (ci, [vp, u64, ctypes.c_char_p, ci]),
        "sun_job_bytes": (ci, [ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_size_t)]),
        "sun_stream_create": (ci, [ctypes.POINTER(vp)]),
        "sun_stream_destroy": (ci, [vp]),
        "sun_stream_synchronize": (ci, [vp]),
        "sun_event_create": (ci, [ctypes.POINTER(vp)]),
        "sun_event_destroy": (ci, [vp]),
        "sun_event_record": (ci, [vp, vp]),
        "sun_event_elapsed_ms": (ci, [vp, vp, ctypes.POINTER(ctypes.c_float)]),
        "sun_device_name": (ci, [vp, ctypes.c_char_p, ci]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().sun_last_error().decode(errors="replace")
        raise SunError(f"{what} failed ({rc}): {msg}", rc)


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """initCULQCD / EndCULQCD (include/comm_mpi.h:43-45)."""

    def __init__(self, device: int = 0, verbosity: int = 0, tune: int = 0):
        self.h = ctypes.c_void_p()
        _check(lib().sun_init(device, verbosity, tune, ctypes.byref(self.h)), "sun_init")

    def close(self):
        if self.h:
            _check(lib().sun_finalize(self.h), "sun_finalize")
            self.h = ctypes.c_void_p()

    def set_launch(self, nc: int, dims, beta: float, storage: int, precision: int, kernel: int, block: int = 256,
                   brick=(-1, 0, 0), lds: int = 0, piped: int = 0):
        """the context's launch shape for lattices of this description (later lattices
        and job-server jobs of it start with it)"""
        d = LatticeDesc(nc, (ctypes.c_int * 4)(*[int(x) for x in dims]), beta, storage, precision)
        c = Launch(block, (ctypes.c_int * 3)(*brick), lds, int(piped))
        _check(lib().sun_ctx_set_launch(self.h, ctypes.byref(d), kernel, ctypes.byref(c)), "sun_ctx_set_launch")

    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(lib().sun_device_name(self.h, buf, 256), "sun_device_name")
        return buf.value.decode()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def lattice_bytes(nc, dims, storage=STORE_18, precision=F64, beta=6.0) -> int:
    d = LatticeDesc(nc, (ctypes.c_int * 4)(*dims), beta, storage, precision)
    return int(lib().sun_lattice_bytes(ctypes.byref(d)))


class Lattice:
    """One gauge field resident in HBM (SETPARAMS + _gauge<Real>, per handle)."""

    def __init__(self, ctx: Context, nc: int, dims, beta: float, storage: int = STORE_18,
                 precision: int = F64):
        self.ctx = ctx
        self.nc, self.dims, self.beta = nc, tuple(int(d) for d in dims), float(beta)
        self.storage, self.precision = storage, precision
        self.V = int(np.prod(self.dims))
        d = LatticeDesc(nc, (ctypes.c_int * 4)(*self.dims), beta, storage, precision)
        self.h = ctypes.c_void_p()
        _check(lib().sun_lattice_create(ctx.h, ctypes.byref(d), ctypes.byref(self.h)), "sun_lattice_create")

    def close(self):
        if self.h:
            _check(lib().sun_lattice_destroy(self.h), "sun_lattice_destroy")
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def copy_from(self, other: "Lattice", stream=None):
        """device-to-device copy of other's links (_gauge::Copy)"""
        _check(lib().sun_lattice_copy(self.h, other.h, stream), "sun_lattice_copy")

    def convert_from(self, other: "Lattice", stream=None):
        """this lattice's field := other's, in this lattice's storage and precision"""
        _check(lib().sun_lattice_convert(self.h, other.h, stream), "sun_lattice_convert")

    @property
    def bytes(self) -> int:
        return lattice_bytes(self.nc, self.dims, self.storage, self.precision, self.beta)

    # -- initialisation
    # -- checkpoints
    def save(self, path: str, fmt: int = FILE_NATIVE, save_prec: int = 8, seed: int = 0, next_sweep: int = 0):
        """native: exact bytes + (seed, next_sweep); reference: SaveBin_Gauge layout"""
        _check(lib().sun_lattice_save(self.h, os.fsencode(path), fmt, save_prec, seed, next_sweep),
               "sun_lattice_save")

    def load(self, path: str, fmt: int = FILE_NATIVE, save_prec: int = 8):
        """returns (seed, next_sweep) for a native checkpoint, (None, None) otherwise"""
        seed, sweep = ctypes.c_uint64(0), ctypes.c_uint32(0)
        _check(lib().sun_lattice_load(self.h, os.fsencode(path), fmt, save_prec, ctypes.byref(seed),
                                      ctypes.byref(sweep)), "sun_lattice_load")
        return (int(seed.value), int(sweep.value)) if fmt == FILE_NATIVE else (None, None)

    # -- launch shapes (kernel 0: heatbath, 1: overrelaxation)
    def get_launch(self, kernel: int) -> Launch:
        out = Launch()
        _check(lib().sun_lattice_get_launch(self.h, kernel, ctypes.byref(out)), "sun_lattice_get_launch")
        return out

    def set_launch(self, kernel: int, block: int = 256, brick=(-1, 0, 0), lds: int = 0, piped: int = 0):
        """piped: 0 plain kernel, 1 pipelined by directions, 2 by half-directions"""
        c = Launch(block, (ctypes.c_int * 3)(*brick), lds, int(piped))
        _check(lib().sun_lattice_set_launch(self.h, kernel, ctypes.byref(c)), "sun_lattice_set_launch")

    def cold_start(self, stream=None):
        _check(lib().sun_cold_start(self.h, stream), "sun_cold_start")

    def hot_start(self, seed: int, stream=None):
        _check(lib().sun_hot_start(self.h, seed, stream), "sun_hot_start")

    # -- updates
    def heatbath(self, seed: int, sweep: int, stream=None):
        _check(lib().sun_heatbath(self.h, seed, sweep, stream), "sun_heatbath")

    def overrelax(self, stream=None):
        _check(lib().sun_overrelax(self.h, stream), "sun_overrelax")

    def heatbath_sub(self, seed: int, sweep: int, parity: int, mu: int, stream=None):
        _check(lib().sun_heatbath_sub(self.h, seed, sweep, parity, mu, stream), "sun_heatbath_sub")

    def overrelax_sub(self, parity: int, mu: int, stream=None):
        _check(lib().sun_overrelax_sub(self.h, parity, mu, stream), "sun_overrelax_sub")

    def reunitarize(self, stream=None):
        _check(lib().sun_reunitarize(self.h, stream), "sun_reunitarize")

    # -- observables
    def plaquette(self, stream=None) -> np.ndarray:
        out = np.zeros(2)
        _check(lib().sun_plaquette(self.h, stream, _dp(out)), "sun_plaquette")
        return out

    def polyakov(self, stream=None) -> complex:
        out = np.zeros(2)
        _check(lib().sun_polyakov(self.h, stream, _dp(out)), "sun_polyakov")
        return complex(out[0], out[1])

    def link_monitor(self, stream=None) -> np.ndarray:
        """[mean Tr U / N (re, im), mean det U (re, im), max |U U^+ - 1|_ij, max |det U - 1|]"""
        out = np.zeros(6)
        _check(lib().sun_link_monitor(self.h, stream, _dp(out)), "sun_link_monitor")
        return out

    def link_trace(self, stream=None) -> complex:
        out = np.zeros(2)
        _check(lib().sun_link_trace(self.h, stream, _dp(out)), "sun_link_trace")
        return complex(out[0], out[1])

    def link_det(self, stream=None) -> complex:
        out = np.zeros(2)
        _check(lib().sun_link_det(self.h, stream, _dp(out)), "sun_link_det")
        return complex(out[0], out[1])

    # -- measurement chain (smearing, multihit, Wilson loops)
    def ape_smear(self, w: float, steps: int, nhit: int, tol: float = 0.0, time_dir: bool = False, stream=None):
        _check(lib().sun_ape_smear(self.h, w, steps, nhit, tol, int(time_dir), stream), "sun_ape_smear")

    def stout_smear(self, rho: float, steps: int, time_dir: bool = False, stream=None):
        _check(lib().sun_stout_smear(self.h, rho, steps, int(time_dir), stream), "sun_stout_smear")

    def hyp_smear(self, alphas=(0.75, 0.6, 0.3), steps: int = 1, nhit: int = 10, tol: float = 0.0,
                  time_dir: bool = False, stream=None):
        a1, a2, a3 = alphas
        _check(lib().sun_hyp_smear(self.h, a1, a2, a3, steps, nhit, tol, int(time_dir), stream), "sun_hyp_smear")

    def multihit(self, out: "Lattice", seed: int, nhit: int, stream=None):
        """out (18-real, same N / extents / precision) <- this field with multihit temporal links"""
        _check(lib().sun_multihit(self.h, out.h, seed, nhit, stream), "sun_multihit")

    def multihit_space(self, out: "Lattice", nhit: int, stream=None):
        """out (18-real) <- this field with spatial links averaged over nhit overrelaxation steps"""
        _check(lib().sun_multihit_space(self.h, out.h, nhit, stream), "sun_multihit_space")

    def multihit_ext(self, out: "Lattice", seed: int, nhit: int, stream=None):
        """out (18-real) <- this field with temporal links averaged over nhit heatbath sweeps of each site's 3^3
        box of temporal links (multihitext.cu)"""
        _check(lib().sun_multihit_ext(self.h, out.h, seed, nhit, stream), "sun_multihit_ext")

    def wilson_loops(self, rmax: int, tmax: int, stream=None) -> np.ndarray:
        """complex [rmax + 1, tmax + 1]: W(r, T)"""
        out = np.zeros((rmax + 1) * (tmax + 1) * 2)
        _check(lib().sun_wilson_loops(self.h, rmax, tmax, _dp(out), stream), "sun_wilson_loops")
        return (out[0::2] + 1j * out[1::2]).reshape(rmax + 1, tmax + 1)

    def wilson_loops_noalloc(self, rmax: int, tmax: int, dev_out: int, stream=None):
        """W(r, T) into DEVICE memory at address dev_out (2 (rmax+1)(tmax+1) doubles,
        16-byte aligned), asynchronously on stream, without allocating: the line field
        and partial sums use the device's fixed scratch array"""
        _check(lib().sun_wilson_loops_noalloc(self.h, rmax, tmax, ctypes.c_void_p(dev_out), stream),
               "sun_wilson_loops_noalloc")

    def wilson_loops_ops(self, rmax: int, tmax: int, family: int = 0, stream=None) -> np.ndarray:
        """[rmax - 1, nop^2, tmax + 1]: for r = 2..rmax the operator-basis correlators of
        every slot pair summed over the spatial directions, / (3 N V) (test.cpp:273-298)"""
        nop = WLOP_NOPS[family]
        out = np.zeros((max(rmax - 1, 0), nop * nop, tmax + 1))
        _check(lib().sun_wilson_loops_ops(self.h, family, rmax, tmax, _dp(out), stream), "sun_wilson_loops_ops")
        return out

    # -- host exchange (canonical layout, complex128 [4, V, N, N])
    def to_host(self) -> np.ndarray:
        U = np.zeros((4, self.V, self.nc, self.nc), dtype=np.complex128)
        _check(lib().sun_links_to_host(self.h, _dp(U), 0), "sun_links_to_host")
        return U

    def from_host(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        assert U.shape == (4, self.V, self.nc, self.nc)
        _check(lib().sun_links_from_host(self.h, _dp(U), 0), "sun_links_from_host")

    # -- generation driver
    def generate(self, seed: int, n_iter: int, n_or: int = 4, reunit_every: int = 1, first_sweep: int = 0,
                 measure: bool = True, use_graph: bool = True, stream=None) -> np.ndarray:
        d = GenerateDesc(seed, first_sweep, n_iter, n_or, reunit_every, int(measure), int(use_graph))
        obs = np.zeros((max(n_iter, 1), 4))
        _check(lib().sun_generate(self.h, ctypes.byref(d), stream, _dp(obs) if measure else None),
               "sun_generate")
        return obs[:n_iter]


WLOP_A0, WLOP_33 = 0, 1
WLOP_NOPS = {WLOP_A0: 10, WLOP_33: 25}


class WilsonLoopOps:
    """Sigma_g_plus (wloopex.h:13-36) / the fieldOp + wloop pair of wlop.h: an
    operator field of one family on the lattice's device and an FP64 accumulator of
    the correlators of every slot pair."""

    def __init__(self, lat: Lattice, family: int, tmax: int):
        self.lat, self.family, self.tmax = lat, family, int(tmax)
        self.h = ctypes.c_void_p()
        _check(lib().sun_wlop_create(lat.h, family, tmax, ctypes.byref(self.h)), "sun_wlop_create")
        self.nop = lib().sun_wlop_nops(self.h)

    def close(self):
        if self.h:
            _check(lib().sun_wlop_destroy(self.h), "sun_wlop_destroy")
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def fields(self, r: int, mu: int, stream=None):
        """CalcWLOPs_A0 / CalcWLOPs_dg_33: the operator field of direction mu at radius r"""
        _check(lib().sun_wlop_fields(self.h, r, mu, stream), "sun_wlop_fields")

    def correlate(self, r: int, mu: int, stream=None):
        """CalcWilsonLoop_A0 / CalcWilsonLoop_dg: add every slot pair's correlators"""
        _check(lib().sun_wlop_correlate(self.h, r, mu, stream), "sun_wlop_correlate")

    def clear(self, stream=None):
        _check(lib().sun_wlop_clear(self.h, stream), "sun_wlop_clear")

    def read(self, stream=None) -> np.ndarray:
        """the accumulator [nop^2, tmax + 1], row i nop + j"""
        out = np.zeros((self.nop * self.nop, self.tmax + 1))
        _check(lib().sun_wlop_read(self.h, _dp(out), stream), "sun_wlop_read")
        return out

    def read_blocks(self, comps, pos, stream=None) -> np.ndarray:
        """the pairs of each component block (wilsonloop_dg.cu:101-122 order)"""
        comps = [int(c) for c in comps]
        pos = [int(p) for p in pos]
        n = sum(c * c for c in comps)
        out = np.zeros((n, self.tmax + 1))
        ca = (ctypes.c_int * max(len(comps), 1))(*comps)
        pa = (ctypes.c_int * max(len(pos), 1))(*pos)
        _check(lib().sun_wlop_read_blocks(self.h, len(comps), ca, pa, _dp(out), stream), "sun_wlop_read_blocks")
        return out

    def fields_to_host(self) -> np.ndarray:
        """complex128 [nop, V, N, N] in the canonical (even/odd) site order"""
        nc = self.lat.nc
        out = np.zeros((self.nop, self.lat.V, nc, nc), dtype=np.complex128)
        _check(lib().sun_wlop_fields_to_host(self.h, _dp(out)), "sun_wlop_fields_to_host")
        return out


class Batch:
    """Many lattices of one (N, storage, precision), one launch per (parity, mu)."""

    def __init__(self, ctx: Context, lattices, seeds):
        self.ctx = ctx
        self.lattices = list(lattices)
        n = len(self.lattices)
        arr = (ctypes.c_void_p * n)(*[l.h.value for l in self.lattices])
        sd = (ctypes.c_uint64 * n)(*[int(s) for s in seeds])
        self.h = ctypes.c_void_p()
        _check(lib().sun_batch_create(ctx.h, n, arr, sd, ctypes.byref(self.h)), "sun_batch_create")

    def close(self):
        if self.h:
            _check(lib().sun_batch_destroy(self.h), "sun_batch_destroy")
            self.h = ctypes.c_void_p()

    def get_launch(self, kernel: int):
        """(piped, brick, lds) of the heatbath (kernel 0) or overrelaxation (1)
        launches: piped 0 plain, 1 / 2 direction / half-direction software-pipelined
        staple; brick 1 the lattices' update bricks, 0 the linear site order; lds the
        unused dynamic LDS per workgroup (occupancy cap)."""
        p, b, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().sun_batch_get_launch(self.h, kernel, ctypes.byref(p), ctypes.byref(b), ctypes.byref(l)),
               "sun_batch_get_launch")
        return p.value, b.value, l.value

    def set_launch(self, kernel: int, piped: int, brick: int = 0, lds: int = 0):
        _check(lib().sun_batch_set_launch(self.h, kernel, piped, brick, lds), "sun_batch_set_launch")

    def heatbath(self, sweeps, stream=None):
        n = len(self.lattices)
        sw = (ctypes.c_uint32 * n)(*[int(s) for s in sweeps])
        _check(lib().sun_batch_heatbath(self.h, sw, stream), "sun_batch_heatbath")

    def overrelax(self, stream=None):
        _check(lib().sun_batch_overrelax(self.h, stream), "sun_batch_overrelax")

    def reunitarize(self, stream=None):
        _check(lib().sun_batch_reunitarize(self.h, stream), "sun_batch_reunitarize")

    def plaquette(self, stream=None) -> np.ndarray:
        out = np.zeros((len(self.lattices), 2))
        _check(lib().sun_batch_plaquette(self.h, stream, _dp(out)), "sun_batch_plaquette")
        return out

    def generate(self, first_sweep: int, n_iter: int, n_or: int = 4, reunit_every: int = 1,
                 use_graph: bool = True, measure: bool = True, stream=None) -> np.ndarray:
        n = len(self.lattices)
        obs = np.zeros((max(n_iter, 1), n, 2))
        _check(lib().sun_batch_generate(self.h, first_sweep, n_iter, n_or, reunit_every, int(use_graph), stream,
                                        _dp(obs) if measure else None), "sun_batch_generate")
        return obs[:n_iter]


def job_bytes(nc: int, dims, beta: float, n_iter: int, n_or: int = 4, reunit_every: int = 1,
              storage: int = STORE_18, precision: int = F64, hot_start: bool = False, seed: int = 0) -> int:
    """server memory a job occupies while resident (sun_job_bytes); raises SunError
    for a description the server refuses.  Host only: needs no device."""
    d = JobDesc(nc, (ctypes.c_int * 4)(*dims), beta, storage, precision, seed, int(hot_start), n_iter, n_or,
                reunit_every)
    b = ctypes.c_size_t()
    _check(lib().sun_job_bytes(ctypes.byref(d), ctypes.byref(b)), "sun_job_bytes")
    return int(b.value)


class Server:
    """Continuous-batching job server (sun_server_*): jobs are admitted while their
    memory fits the HBM budget reserved at creation; each scheduling step runs one
    iteration of every resident job."""

    def __init__(self, ctx: Context, budget_bytes: int, max_jobs: int = 256):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        _check(lib().sun_server_create(ctx.h, int(budget_bytes), int(max_jobs), ctypes.byref(self.h)),
               "sun_server_create")

    def close(self):
        if self.h:
            _check(lib().sun_server_destroy(self.h), "sun_server_destroy")
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, nc: int, dims, beta: float, seed: int, n_iter: int, n_or: int = 4, reunit_every: int = 1,
               storage: int = STORE_18, precision: int = F64, hot_start: bool = False) -> int:
        d = JobDesc(nc, (ctypes.c_int * 4)(*dims), beta, storage, precision, seed, int(hot_start), n_iter, n_or,
                    reunit_every)
        jid = ctypes.c_uint64()
        _check(lib().sun_server_submit(self.h, ctypes.byref(d), ctypes.byref(jid)), "sun_server_submit")
        return int(jid.value)

    def cancel(self, jid: int):
        _check(lib().sun_server_cancel(self.h, jid), "sun_server_cancel")

    def step(self):
        _check(lib().sun_server_step(self.h), "sun_server_step")

    def start(self):
        _check(lib().sun_server_start(self.h), "sun_server_start")

    def stop(self):
        _check(lib().sun_server_stop(self.h), "sun_server_stop")

    def wait(self, jid: int, timeout_s: float = 60.0):
        _check(lib().sun_server_wait(self.h, jid, timeout_s), "sun_server_wait")

    def status(self, jid: int) -> JobStatus:
        st = JobStatus()
        _check(lib().sun_server_status(self.h, jid, ctypes.byref(st)), "sun_server_status")
        return st

    def observables(self, jid: int, n_iter: int) -> np.ndarray:
        """rows (plaq_spatial, plaq_temporal, poly_re, poly_im), one per iteration"""
        out = np.zeros((max(n_iter, 1), 4))
        _check(lib().sun_server_observables(self.h, jid, _dp(out), n_iter), "sun_server_observables")
        return out[:n_iter]

    def release(self, jid: int):
        _check(lib().sun_server_release(self.h, jid), "sun_server_release")

    def stats(self) -> ServerCounters:
        st = ServerCounters()
        _check(lib().sun_server_stats(self.h, ctypes.byref(st)), "sun_server_stats")
        return st

    def clock(self) -> float:
        return float(lib().sun_server_clock(self.h))

    def suspend(self, jid: int, path: str):
        """take the job off the device into a file (a running job: at its next
        iteration boundary; it then reaches JOB_SUSPENDED)"""
        _check(lib().sun_server_suspend(self.h, jid, os.fsencode(path)), "sun_server_suspend")

    def resume(self, path: str) -> int:
        """submit a suspended job again; returns its new id"""
        jid = ctypes.c_uint64()
        _check(lib().sun_server_resume(self.h, os.fsencode(path), ctypes.byref(jid)), "sun_server_resume")
        return int(jid.value)

    def job_error(self, jid: int) -> str:
        """why a job is JOB_FAILED ('' otherwise)"""
        buf = ctypes.create_string_buffer(512)
        _check(lib().sun_server_job_error(self.h, jid, buf, len(buf)), "sun_server_job_error")
        return buf.value.decode()


class Timer:
    """HIP events on a given stream (sun_event_*), for timing without torch."""

    def __init__(self):
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().sun_event_create(ctypes.byref(self.a)), "sun_event_create")
        _check(lib().sun_event_create(ctypes.byref(self.b)), "sun_event_create")

    def start(self, stream=None):
        _check(lib().sun_event_record(self.a, stream), "sun_event_record")

    def stop(self, stream=None):
        _check(lib().sun_event_record(self.b, stream), "sun_event_record")

    def ms(self) -> float:
        v = ctypes.c_float()
        _check(lib().sun_event_elapsed_ms(self.a, self.b, ctypes.byref(v)), "sun_event_elapsed_ms")
        return float(v.value)

    def close(self):
        lib().sun_event_destroy(self.a)
        lib().sun_event_destroy(self.b)


def new_stream():
    s = ctypes.c_void_p()
    _check(lib().sun_stream_create(ctypes.byref(s)), "sun_stream_create")
    return s


def sync(stream):
    _check(lib().sun_stream_synchronize(stream), "sun_stream_synchronize")
