"""Checkpoints: native (exact bytes, resume header, CRC32) and the reference's
SaveBin_Gauge layout (io_gauge.cpp:429-583).

* a native save/load restores every variant bit for bit, and a chain resumed from a
  checkpoint (seed, next sweep) continues exactly as the uninterrupted chain;
* the payload checksum is CRC-32 (checked against zlib) and a corrupted or
  mismatched file is refused with SUN_ERR_INVALID;
* the reference layout is read back by an independent numpy reader of that layout
  (header grid/beta/precision, lexicographic sites, four N x N complex matrices per
  site) and loads back into a lattice, in FP64 and FP32 files, with and without
  header; grid or beta mismatches are refused.
"""
import os
import zlib

import numpy as np
import pytest

import sun_amd
from test_gpu_parity import IDS, VARIANTS

pytestmark = pytest.mark.gpu

HEADER = 80  # native header bytes


def hot(ctx, nc, dims, storage, prec, seed=5, beta=6.0):
    lat = sun_amd.Lattice(ctx, nc, dims, beta, storage, prec)
    lat.hot_start(seed)
    return lat


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
def test_native_roundtrip_bit_exact(ctx, tmp_path, nc, storage, prec):
    dims = (4, 4, 6, 8)
    path = str(tmp_path / "c.sun")
    with hot(ctx, nc, dims, storage, prec) as a, sun_amd.Lattice(ctx, nc, dims, 6.0, storage, prec) as b:
        a.generate(seed=9, n_iter=2, n_or=2)
        a.save(path, seed=9, next_sweep=2)
        seed, sweep = b.load(path)
        assert (seed, sweep) == (9, 2)
        assert np.array_equal(a.to_host(), b.to_host())
        raw = open(path, "rb").read()
        assert len(raw) == HEADER + a.bytes
        crc = int.from_bytes(raw[72:76], "little")  # header offset of payload_crc32
        assert crc == zlib.crc32(raw[HEADER:])


def test_resumed_chain_equals_uninterrupted(ctx, tmp_path):
    dims, seed = (8, 8, 8, 8), 31
    path = str(tmp_path / "chain.sun")
    with hot(ctx, 3, dims, 12, 4, seed) as full:
        obs_full = full.generate(seed=seed, n_iter=6, n_or=4, reunit_every=1)
        U_full = full.to_host()
    with hot(ctx, 3, dims, 12, 4, seed) as first:
        obs_a = first.generate(seed=seed, n_iter=3, n_or=4, reunit_every=1)
        first.save(path, seed=seed, next_sweep=3)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as resumed:
        s, k = resumed.load(path)
        obs_b = resumed.generate(seed=s, n_iter=3, n_or=4, reunit_every=1, first_sweep=k)
        assert np.array_equal(np.concatenate([obs_a, obs_b]), obs_full)
        assert np.array_equal(resumed.to_host(), U_full)


def test_native_refuses_bad_files(ctx, tmp_path):
    dims = (4, 4, 4, 4)
    path = str(tmp_path / "c.sun")
    with hot(ctx, 3, dims, 18, 8) as a:
        a.save(path, seed=1, next_sweep=0)
        raw = bytearray(open(path, "rb").read())
        raw[HEADER + 100] ^= 0x01
        bad = str(tmp_path / "bad.sun")
        open(bad, "wb").write(bytes(raw))
        with pytest.raises(sun_amd.SunError, match="checksum") as e:
            a.load(bad)
        assert e.value.code == sun_amd.ERR_INVALID
        open(bad, "wb").write(bytes(raw[:HEADER + 10]))
        with pytest.raises(sun_amd.SunError, match="truncated"):
            a.load(bad)
        with pytest.raises(sun_amd.SunError):
            a.load(str(tmp_path / "missing.sun"))
    for other in [(3, (4, 4, 4, 6), 18, 8), (3, dims, 12, 8), (3, dims, 18, 4), (2, dims, 18, 8)]:
        nc, d, st, pr = other
        with sun_amd.Lattice(ctx, nc, d, 6.0, st, pr) as b:
            with pytest.raises(sun_amd.SunError, match="does not match") as e:
                b.load(path)
            assert e.value.code == sun_amd.ERR_INVALID


def read_reference(path, nc, dims, header, prec_bytes=None, beta_bytes=8):
    """independent reader of the SaveBin_Gauge layout -> canonical EO array"""
    raw = open(path, "rb").read()
    off = 0
    if header:
        grid = np.frombuffer(raw, dtype=np.int32, count=4)
        off = 16
        beta = np.frombuffer(raw, dtype=np.float64 if beta_bytes == 8 else np.float32, count=1, offset=off)[0]
        off += beta_bytes
        prec_bytes = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=off)[0])
        off += 8
    else:
        grid, beta = np.array(dims), None
    dt = np.float64 if prec_bytes == 8 else np.float32
    V = int(np.prod(dims))
    data = np.frombuffer(raw, dtype=dt, offset=off).astype(np.float64)
    assert data.size == V * 4 * nc * nc * 2
    m = data.reshape(V, 4, nc, nc, 2)
    m = m[..., 0] + 1j * m[..., 1]
    U = np.zeros((4, V, nc, nc), dtype=complex)
    X = dims
    for lex in range(V):
        x0 = lex % X[0]
        x1 = (lex // X[0]) % X[1]
        x2 = (lex // (X[0] * X[1])) % X[2]
        x3 = lex // (X[0] * X[1] * X[2])
        e = lex // 2 + ((x0 + x1 + x2 + x3) & 1) * (V // 2)
        U[:, e] = m[lex]
    return tuple(grid), beta, prec_bytes, U


@pytest.mark.parametrize("nc,storage,prec", [(3, 18, 8), (3, 12, 4), (2, 18, 4), (4, 18, 8), (3, 8, 4)],
                         ids=["su3_s18_f64", "su3_s12_f32", "su2_f32", "su4_f64", "su3_s8_f32"])
@pytest.mark.parametrize("header", [True, False])
def test_reference_layout(ctx, tmp_path, nc, storage, prec, header):
    dims, beta = (4, 6, 4, 8), 5.75
    path = str(tmp_path / "ref.bin")
    fmt = sun_amd.FILE_REFERENCE if header else sun_amd.FILE_REFERENCE_NOHEADER
    with hot(ctx, nc, dims, storage, prec, beta=beta) as a, sun_amd.Lattice(ctx, nc, dims, beta, storage, prec) as b:
        G = a.to_host()
        for save_prec in (8, 4):
            a.save(path, fmt, save_prec=save_prec)
            grid, fbeta, fprec, U = read_reference(path, nc, dims, header, save_prec, beta_bytes=prec)
            assert grid == dims and fprec == save_prec
            if header:
                assert fbeta == (np.float32(beta) if prec == 4 else beta)
            tol = 0.0 if save_prec == 8 else 1e-7
            assert np.abs(U - G).max() <= tol
            b.load(path, fmt, save_prec=save_prec)
            # exact when the file holds the stored values exactly (FP64 file, or an
            # FP32 lattice) and the storage keeps them (18/12 reals); 8-real storage
            # re-derives its phases (atan2), an FP32 file of an FP64 lattice rounds
            H = b.to_host()
            if storage == 8:
                lim = 2e-5
            elif save_prec == 8 or prec == 4:
                lim = 0.0
            else:
                lim = 1e-6
            assert np.abs(H - G).max() <= lim, (save_prec, np.abs(H - G).max())


def test_reference_header_mismatch(ctx, tmp_path):
    path = str(tmp_path / "ref.bin")
    with hot(ctx, 3, (4, 4, 4, 4), 18, 8, beta=6.0) as a:
        a.save(path, sun_amd.FILE_REFERENCE)
    with sun_amd.Lattice(ctx, 3, (4, 4, 4, 4), 6.2, 18, 8) as b:
        with pytest.raises(sun_amd.SunError, match="beta") as e:
            b.load(path, sun_amd.FILE_REFERENCE)
        assert e.value.code == sun_amd.ERR_INVALID
    with sun_amd.Lattice(ctx, 3, (4, 4, 4, 6), 6.0, 18, 8) as b:
        with pytest.raises(sun_amd.SunError, match="grid"):
            b.load(path, sun_amd.FILE_REFERENCE)
    with sun_amd.Lattice(ctx, 2, (4, 4, 4, 4), 6.0, 18, 8) as b:
        with pytest.raises(sun_amd.SunError):
            b.load(path, sun_amd.FILE_REFERENCE)
    assert os.path.getsize(path) == 16 + 8 + 8 + 256 * 4 * 9 * 16
