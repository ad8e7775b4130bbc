"""ctypes binding of the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP engine is compared
with; it is imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` and by nothing on the product path.

Fields are numpy ``complex128`` arrays of shape ``(4, V, N, N)``: direction, then
even/odd site index ``eo(x) = parity(x)*V/2 + lex(x)/2`` (include/index.h:133-141
of the reference), then the row-major colour matrix.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

STORE_18, STORE_12, STORE_8 = 18, 12, 8
TAG_HB, TAG_HOT, TAG_TEST = 0x10000000, 0x20000000, 0x30000000


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64, u32, ci, cd = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
    sig = {
        "or_philox4x32_10": (None, [u32p, u32p, u32p]),
        "or_u01": (cd, [u32]),
        "or_stream_uniforms": (None, [u64, u32, u32, u32, ci, dp]),
        "or_generate_su2_milc": (ci, [cd, dp, ci, dp]),
        "or_sample_a0": (None, [cd, u64, ci, dp]),
        "or_eo_to_coords": (None, [ci, ci, ip, ip]),
        "or_coords_to_eo": (ci, [ip, ip]),
        "or_neighbor_eo": (ci, [ci, ci, ip, ci, ci, ci, ci]),
        "or_cold_start": (None, [ci, ip, dp]),
        "or_hot_start": (None, [ci, ip, u64, ci, dp]),
        "or_encode12": (None, [dp, dp]),
        "or_decode12": (None, [dp, dp]),
        "or_encode8": (None, [dp, dp]),
        "or_decode8": (None, [dp, dp]),
        "or_project_storage": (None, [ci, ci, ctypes.c_long, dp]),
        "or_staple": (None, [ci, ip, dp, ci, ci, ci, dp]),
        "or_heatbath_sub": (ci, [ci, ip, cd, u64, u32, ci, dp, ci, ci]),
        "or_overrelax_sub": (ci, [ci, ip, ci, dp, ci, ci]),
        "or_heatbath_sweep": (ci, [ci, ip, cd, u64, u32, ci, dp]),
        "or_heatbath_sweep_margins": (ci, [ci, ip, cd, u64, u32, ci, dp, dp]),
        "or_overrelax_sweep": (ci, [ci, ip, ci, dp]),
        "or_overrelax_link": (None, [ci, dp, dp]),
        "or_heatbath_link": (None, [ci, cd, dp, dp, u64, u32, u32, u32]),
        "or_reunit_link": (None, [ci, dp]),
        "or_reunitarize": (None, [ci, ctypes.c_long, ci, dp]),
        "or_plaquette": (None, [ci, ip, dp, dp]),
        "or_polyakov": (None, [ci, ip, dp, dp]),
        "or_link_monitor": (None, [ci, ctypes.c_long, dp, dp]),
        "or_project_link": (None, [ci, dp, dp, ci, cd]),
        "or_su3_project": (None, [dp, ci]),
        "or_ape_smear": (ci, [ci, ip, dp, cd, ci, ci, cd, ci]),
        "or_multihit": (ci, [ci, ip, cd, u64, ci, dp, dp]),
        "or_multihit_margins": (ci, [ci, ip, cd, u64, ci, dp, dp, dp]),
        "or_multihit_ext_margins": (ci, [ci, ip, cd, u64, ci, dp, dp, dp]),
        "or_exp_iq": (None, [dp, dp]),
        "or_stout_smear": (ci, [ip, dp, cd, ci, ci]),
        "or_multihit_space": (ci, [ci, ip, ci, dp, dp]),
        "or_multihit_ext": (ci, [ci, ip, cd, u64, ci, dp, dp]),
        "or_hyp_smear": (ci, [ci, ip, dp, cd, cd, cd, ci, ci, cd, ci]),
        "or_wilson_loops": (ci, [ci, ip, dp, ci, ci, dp]),
        "or_wlop_fields": (ci, [ci, ip, dp, ci, ci, ci, dp]),
        "or_wlop_correlate": (ci, [ci, ip, dp, dp, ci, ci, ci, ci, dp]),
        "or_matmul": (None, [ci, dp, dp, dp]),
        "or_det": (None, [ci, dp, dp]),
        "or_num_threads": (ci, []),
        "or_set_num_threads": (None, [ci]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dims(dims):
    return (ctypes.c_int * 4)(*[int(d) for d in dims])


def volume(dims) -> int:
    return int(np.prod(dims))


def empty_field(nc: int, dims) -> np.ndarray:
    return np.zeros((4, volume(dims), nc, nc), dtype=np.complex128)


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"oracle {what} rejected its arguments (code {rc})")


# ---- RNG -----------------------------------------------------------------------
def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return tuple(int(x) for x in o)


def u01(w: int) -> float:
    """uniform in (0,1] from one 32-bit word: (w + 1) * 2^-32"""
    return lib().or_u01(w)


def stream_uniforms(seed: int, c0: int, c2: int, c3: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().or_stream_uniforms(seed, c0, c2, c3, n, _dp(out))
    return out


def generate_su2(al: float, uniforms) -> tuple[np.ndarray, int]:
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    a = np.zeros(4)
    used = lib().or_generate_su2_milc(al, _dp(u), len(u), _dp(a))
    return a, used


def sample_a0(al: float, seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().or_sample_a0(al, seed, n, _dp(out))
    return out


# ---- geometry ------------------------------------------------------------------
def eo_to_coords(idx: int, parity: int, dims):
    x = (ctypes.c_int * 4)()
    lib().or_eo_to_coords(idx, parity, _dims(dims), x)
    return tuple(x)


def coords_to_eo(x, dims) -> int:
    return lib().or_coords_to_eo((ctypes.c_int * 4)(*x), _dims(dims))


def neighbor_eo(idx, parity, dims, mu, lmu, nu=-1, lnu=0) -> int:
    return lib().or_neighbor_eo(idx, parity, _dims(dims), mu, lmu, nu, lnu)


# ---- fields ----------------------------------------------------------------------
def cold_start(nc: int, dims) -> np.ndarray:
    U = empty_field(nc, dims)
    lib().or_cold_start(nc, _dims(dims), _dp(U))
    return U


def hot_start(nc: int, dims, seed: int, storage: int = STORE_18) -> np.ndarray:
    U = empty_field(nc, dims)
    lib().or_hot_start(nc, _dims(dims), seed, storage, _dp(U))
    return U


def project_storage(U: np.ndarray, storage: int) -> np.ndarray:
    U = np.ascontiguousarray(U.copy())
    nc = U.shape[-1]
    lib().or_project_storage(nc, storage, U.size // (nc * nc), _dp(U))
    return U


def encode12(link):
    link = np.ascontiguousarray(link, dtype=np.complex128)
    out = np.zeros(12)
    lib().or_encode12(_dp(link), _dp(out))
    return out


def decode12(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros((3, 3), dtype=np.complex128)
    lib().or_decode12(_dp(p), _dp(out))
    return out


def encode8(link):
    link = np.ascontiguousarray(link, dtype=np.complex128)
    out = np.zeros(8)
    lib().or_encode8(_dp(link), _dp(out))
    return out


def decode8(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros((3, 3), dtype=np.complex128)
    lib().or_decode8(_dp(p), _dp(out))
    return out


def staple(U: np.ndarray, dims, parity: int, mu: int, idx: int) -> np.ndarray:
    nc = U.shape[-1]
    S = np.zeros((nc, nc), dtype=np.complex128)
    lib().or_staple(nc, _dims(dims), _dp(U), parity, mu, idx, _dp(S))
    return S


def heatbath_sub(U, dims, beta, seed, sweep, parity, mu, storage=STORE_18):
    nc = U.shape[-1]
    _check(lib().or_heatbath_sub(nc, _dims(dims), beta, seed, sweep, storage, _dp(U), parity, mu), "heatbath")
    return U


def overrelax_sub(U, dims, parity, mu, storage=STORE_18):
    nc = U.shape[-1]
    _check(lib().or_overrelax_sub(nc, _dims(dims), storage, _dp(U), parity, mu), "overrelax")
    return U


def heatbath_sweep(U, dims, beta, seed, sweep, storage=STORE_18):
    nc = U.shape[-1]
    _check(lib().or_heatbath_sweep(nc, _dims(dims), beta, seed, sweep, storage, _dp(U)), "heatbath")
    return U


def heatbath_sweep_margins(U, dims, beta, seed, sweep, storage=STORE_18):
    """heatbath sweep of U in place; returns, per link (shape (4, V)), the smallest
    distance from a tie among the sampler decisions of its update (inf if none)"""
    nc = U.shape[-1]
    m = np.empty(U.shape[:2], dtype=np.float64)
    _check(lib().or_heatbath_sweep_margins(nc, _dims(dims), beta, seed, sweep, storage, _dp(U), _dp(m)),
           "heatbath")
    return m


def overrelax_sweep(U, dims, storage=STORE_18):
    nc = U.shape[-1]
    _check(lib().or_overrelax_sweep(nc, _dims(dims), storage, _dp(U)), "overrelax")
    return U


def overrelax_link(U, F):
    U = np.ascontiguousarray(U, dtype=np.complex128).copy()
    F = np.ascontiguousarray(F, dtype=np.complex128)
    lib().or_overrelax_link(U.shape[-1], _dp(U), _dp(F))
    return U


def heatbath_link(U, F, beta, seed, c0=0, c2=0, c3=TAG_HB):
    U = np.ascontiguousarray(U, dtype=np.complex128).copy()
    F = np.ascontiguousarray(F, dtype=np.complex128)
    lib().or_heatbath_link(U.shape[-1], beta, _dp(U), _dp(F), seed, c0, c2, c3)
    return U


def reunit_link(link):
    link = np.ascontiguousarray(link, dtype=np.complex128).copy()
    lib().or_reunit_link(link.shape[-1], _dp(link))
    return link


def reunitarize(U, storage=STORE_18):
    nc = U.shape[-1]
    lib().or_reunitarize(nc, U.size // (nc * nc), storage, _dp(U))
    return U


def plaquette(U, dims):
    out = np.zeros(2)
    lib().or_plaquette(U.shape[-1], _dims(dims), _dp(U), _dp(out))
    return out


def polyakov(U, dims):
    out = np.zeros(2)
    lib().or_polyakov(U.shape[-1], _dims(dims), _dp(U), _dp(out))
    return complex(out[0], out[1])


def link_monitor(U) -> np.ndarray:
    """[mean Tr U / N re, im, mean det U re, im, max |U U^+ - 1|_ij, max |det U - 1|]"""
    nc = U.shape[-1]
    out = np.zeros(6)
    lib().or_link_monitor(nc, U.size // (4 * nc * nc), _dp(np.ascontiguousarray(U)), _dp(out))
    return out


def project_link(w, q, nhit: int, tol: float = 0.0):
    w = np.array(w, dtype=np.complex128)
    q = np.ascontiguousarray(q, dtype=np.complex128)
    lib().or_project_link(w.shape[-1], _dp(w), _dp(q), nhit, tol)
    return w


def su3_project(F, nhit: int):
    F = np.array(F, dtype=np.complex128)
    lib().or_su3_project(_dp(F), nhit)
    return F


def ape_smear(U, dims, w: float, steps: int, nhit: int, tol: float = 0.0, time_dir: bool = False):
    """in place on an 18-real field"""
    _check(lib().or_ape_smear(U.shape[-1], _dims(dims), _dp(U), w, steps, nhit, tol, int(time_dir)), "ape_smear")
    return U


def exp_iq(Q):
    Q = np.ascontiguousarray(Q, dtype=np.complex128)
    E = np.zeros_like(Q)
    lib().or_exp_iq(_dp(Q), _dp(E))
    return E


def stout_smear(U, dims, rho: float, steps: int, time_dir: bool = False):
    """in place on an SU(3) 18-real field"""
    _check(lib().or_stout_smear(_dims(dims), _dp(U), rho, steps, int(time_dir)), "stout_smear")
    return U


def hyp_smear(U, dims, alphas=(0.75, 0.6, 0.3), steps: int = 1, nhit: int = 10, tol: float = 0.0,
              time_dir: bool = False):
    """in place on an 18-real field"""
    a1, a2, a3 = alphas
    _check(lib().or_hyp_smear(U.shape[-1], _dims(dims), _dp(U), a1, a2, a3, steps, nhit, tol, int(time_dir)),
           "hyp_smear")
    return U


def multihit(U, dims, beta: float, seed: int, nhit: int):
    out = np.empty_like(U)
    _check(lib().or_multihit(U.shape[-1], _dims(dims), beta, seed, nhit, _dp(U), _dp(out)), "multihit")
    return out


def multihit_ext(U, dims, beta: float, seed: int, nhit: int):
    out = np.empty_like(U)
    _check(lib().or_multihit_ext(U.shape[-1], _dims(dims), beta, seed, nhit, _dp(U), _dp(out)), "multihit_ext")
    return out


def multihit_margins(U, dims, beta: float, seed: int, nhit: int):
    """(multihit result, per-link smallest distance from a sampler tie; temporal links)"""
    out = np.empty_like(U)
    m = np.full(U.shape[:2], np.inf)
    _check(lib().or_multihit_margins(U.shape[-1], _dims(dims), beta, seed, nhit, _dp(U), _dp(out), _dp(m)),
           "multihit")
    return out, m


def multihit_ext_margins(U, dims, beta: float, seed: int, nhit: int):
    """(extended multihit result, per temporal link the smallest distance from a
    sampler tie over every update of its box)"""
    out = np.empty_like(U)
    m = np.full(U.shape[:2], np.inf)
    _check(lib().or_multihit_ext_margins(U.shape[-1], _dims(dims), beta, seed, nhit, _dp(U), _dp(out), _dp(m)),
           "multihit_ext")
    return out, m


def multihit_space(U, dims, nhit: int):
    out = np.empty_like(U)
    _check(lib().or_multihit_space(U.shape[-1], _dims(dims), nhit, _dp(U), _dp(out)), "multihit_space")
    return out


def wilson_loops(U, dims, rmax: int, tmax: int) -> np.ndarray:
    """complex [rmax + 1, tmax + 1]"""
    out = np.zeros((rmax + 1) * (tmax + 1) * 2)
    _check(lib().or_wilson_loops(U.shape[-1], _dims(dims), _dp(U), rmax, tmax, _dp(out)), "wilson_loops")
    return (out[0::2] + 1j * out[1::2]).reshape(rmax + 1, tmax + 1)


# ---- operator-basis Wilson loops (src/wl/*.cu) ------------------------------------
WLOP_A0, WLOP_33 = 0, 1
WLOP_NOPS = {WLOP_A0: 10, WLOP_33: 25}


def wlop_fields(U, dims, family: int, r: int, mu: int) -> np.ndarray:
    """operator fields [nop, V, N, N] of direction mu at radius r (calcop_dg_A0.cu,
    calcop_dg_33.cu)"""
    nc = U.shape[-1]
    ops = np.zeros((WLOP_NOPS[family], volume(dims), nc, nc), dtype=np.complex128)
    _check(lib().or_wlop_fields(nc, _dims(dims), _dp(U), family, r, mu, _dp(ops)), "wlop_fields")
    return ops


def wlop_correlate(U, dims, ops, r: int, mu: int, tmax: int, out=None) -> np.ndarray:
    """adds sum_x Re Tr(O_i(x) t1 O_j(x + T e3)^+ t0^+) to out[i nop + j, T]
    (wilsonloop_dg_A0.cu:57-105) and returns out"""
    nop = ops.shape[0]
    if out is None:
        out = np.zeros((nop * nop, tmax + 1))
    assert out.shape == (nop * nop, tmax + 1) and out.flags.c_contiguous
    ops = np.ascontiguousarray(ops)
    _check(lib().or_wlop_correlate(U.shape[-1], _dims(dims), _dp(U), _dp(ops), nop, r, mu, tmax, _dp(out)),
           "wlop_correlate")
    return out


def wilson_loops_ops(U, dims, rmax: int, tmax: int, family: int = WLOP_A0) -> np.ndarray:
    """test.cpp:273-298: for r = 2..rmax the sum over mu < 3 of the correlators of
    every operator pair, / (3 N V); [rmax - 1, nop^2, tmax + 1]"""
    nc = U.shape[-1]
    nop = WLOP_NOPS[family]
    res = np.zeros((max(rmax - 1, 0), nop * nop, tmax + 1))
    for r in range(2, rmax + 1):
        for mu in range(3):
            wlop_correlate(U, dims, wlop_fields(U, dims, family, r, mu), r, mu, tmax, res[r - 2])
        res[r - 2] /= volume(dims) * 3 * nc
    return res


def matmul(A, B):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    C = np.zeros_like(A)
    lib().or_matmul(A.shape[-1], _dp(A), _dp(B), _dp(C))
    return C


def det(A) -> complex:
    A = np.ascontiguousarray(A, dtype=np.complex128)
    out = np.zeros(2)
    lib().or_det(A.shape[-1], _dp(A), _dp(out))
    return complex(out[0], out[1])


def set_threads(n: int):
    lib().or_set_num_threads(n)


def threads() -> int:
    return lib().or_num_threads()
