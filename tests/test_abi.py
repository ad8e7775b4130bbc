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
