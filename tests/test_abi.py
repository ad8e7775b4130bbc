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
