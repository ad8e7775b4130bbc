"""The C-ABI library loads and exports every symbol include/sun.h declares (CPU:
no compute calls are made)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    hdr = open(os.path.join(ROOT, "include", "sun.h")).read()
    return sorted(set(re.findall(r"^[a-z_ ]+\*?\s*(sun_[a-z0-9_]+)\(", hdr, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["sun_init", "sun_finalize", "sun_lattice_create", "sun_heatbath", "sun_overrelax",
                 "sun_reunitarize", "sun_plaquette", "sun_polyakov", "sun_generate", "sun_batch_generate"]:
        assert must in names


def test_library_exports_all_declared_symbols():
    import sun_amd
    L = sun_amd.lib()
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", sun_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (sun_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared()) <= exported
    assert set(sun_amd.EXPORTS) == set(declared())


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import sun_amd
    monkeypatch.setattr(sun_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(sun_amd, "_lib", None)
    with pytest.raises(sun_amd.SunError, match="missing"):
        sun_amd.lib()


def test_facade_templates_compile(tmp_path):
    """every class and function template of include/sun.hpp instantiates (both
    precisions) and links against libsun_hip.so"""
    import sun_amd
    src = os.path.join(ROOT, "tests", "host", "facade_all.cpp")
    obj = str(tmp_path / "facade_all.so")
    r = subprocess.run(["g++", "-std=c++17", "-shared", "-fPIC", "-Wall", "-Werror", f"-I{ROOT}/include", src,
                        "-o", obj, f"-L{os.path.dirname(sun_amd.LIB_PATH)}", "-lsun_hip"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_descriptor_limits():
    """the largest accepted volume (4V x 16-byte records addressable by 32-bit plane
    offsets: V < 2^26, e.g. 90^4) and the rejected shapes; host-only calls"""
    import sun_amd
    assert sun_amd.lattice_bytes(3, (90, 90, 90, 90), 12, 4) == 4 * 90 ** 4 * 48
    assert sun_amd.lattice_bytes(3, (92, 92, 92, 92), 12, 4) == 0       # 4V x 16 B > 2^32
    assert sun_amd.lattice_bytes(3, (2, 2, 2, 2), 18, 8) == 4 * 16 * 144  # the smallest lattice
    for bad in [(3, (4, 4, 4, 5), 18, 8), (3, (0, 4, 4, 4), 18, 8), (5, (4, 4, 4, 4), 18, 8),
                (2, (4, 4, 4, 4), 12, 4), (3, (4, 4, 4, 4), 10, 4), (3, (4, 4, 4, 4), 18, 2)]:
        assert sun_amd.lattice_bytes(*bad) == 0, bad
    # extents whose product overflows 64 bits are refused before the product wraps
    # (65536^4 = 2^64 would wrap to 0 and pass a check made after the loop)
    assert sun_amd.lattice_bytes(3, (65536, 65536, 65536, 65536), 18, 8) == 0
    assert sun_amd.lattice_bytes(3, (2 ** 30, 2 ** 30, 2, 2), 18, 8) == 0


def test_job_validation_on_host():
    """sun_job_bytes applies the server's job checks without a device: the memory a
    job occupies, and the refusals (12-real links without reunitarisation every
    iteration, cold 8-real starts, reunit_every > 1, bad extents and counts)"""
    import sun_amd
    b = sun_amd.job_bytes(3, (8, 8, 8, 8), 6.0, 50, storage=12, precision=4)
    assert b >= sun_amd.lattice_bytes(3, (8, 8, 8, 8), 12, 4) and b % 256 == 0
    # the observable history grows with the iterations
    assert sun_amd.job_bytes(3, (8, 8, 8, 8), 6.0, 500, storage=12, precision=4) > b
    assert sun_amd.job_bytes(3, (8, 8, 8, 8), 6.0, 50, storage=18, precision=4, reunit_every=0) > 0
    assert sun_amd.job_bytes(3, (8, 8, 8, 8), 6.0, 50, storage=8, precision=4, hot_start=True, reunit_every=0) > 0
    cases = [
        (dict(storage=12, precision=4, reunit_every=0), sun_amd.ERR_UNSUPPORTED),
        (dict(storage=12, precision=8, reunit_every=0), sun_amd.ERR_UNSUPPORTED),
        (dict(storage=8, precision=4, hot_start=False), sun_amd.ERR_UNSUPPORTED),
        (dict(reunit_every=2), sun_amd.ERR_UNSUPPORTED),
        (dict(n_iter=0), sun_amd.ERR_INVALID),
        (dict(n_or=-1), sun_amd.ERR_INVALID),
        (dict(dims=(8, 8, 8, 7)), sun_amd.ERR_INVALID),
        (dict(nc=5), sun_amd.ERR_UNSUPPORTED),
    ]
    for kw, code in cases:
        args = dict(nc=3, dims=(8, 8, 8, 8), beta=6.0, n_iter=10)
        args.update(kw)
        with pytest.raises(sun_amd.SunError) as e:
            sun_amd.job_bytes(**args)
        assert e.value.code == code, (kw, e.value.code, str(e.value))
