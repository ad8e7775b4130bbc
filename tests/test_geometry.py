"""Even/odd geometry (reference include/index.h:133-202), CPU only.

Pins the oracle's site numbering by brute force (an independent numpy
enumeration of eo(x) = parity(x)*V/2 + lex(x)/2) and checks the engine's device
geometry (sun_field.hpp: FastDiv, site_coords, lex_shift), compiled for the host,
against the same enumeration -- including that every neighbour index lies in
[0, V), so no kernel reads out of bounds.  Also checks the schedule's race freedom:
in a (parity, mu) launch no site reads a link another site writes.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DIMS = [(2, 2, 2, 2), (4, 4, 4, 4), (4, 6, 8, 10), (6, 2, 4, 12), (12, 12, 12, 12), (8, 8, 8, 8), (32, 4, 4, 8), (24, 8, 8, 8)]


def brute(dims):
    """coords[eo] and eo_of[x0,x1,x2,x3] by direct enumeration."""
    X = dims
    V = int(np.prod(X))
    eo_of = np.zeros(X[::-1], dtype=np.int64)  # indexed [x3,x2,x1,x0]
    coords = np.zeros((V, 4), dtype=np.int64)
    for x3 in range(X[3]):
        for x2 in range(X[2]):
            for x1 in range(X[1]):
                for x0 in range(X[0]):
                    lex = x0 + X[0] * (x1 + X[1] * (x2 + X[2] * x3))
                    par = (x0 + x1 + x2 + x3) & 1
                    e = par * (V // 2) + lex // 2
                    eo_of[x3, x2, x1, x0] = e
                    coords[e] = (x0, x1, x2, x3)
    return coords, eo_of


def eo_at(eo_of, x, dims):
    y = [int(x[i]) % dims[i] for i in range(4)]
    return int(eo_of[y[3], y[2], y[1], y[0]])


@pytest.mark.parametrize("dims", DIMS[:4])
def test_oracle_eo_bijection_and_neighbours(oracle, dims):
    coords, eo_of = brute(dims)
    V = int(np.prod(dims))
    assert sorted(eo_of.ravel().tolist()) == list(range(V))
    for e in range(V):
        parity, idx = divmod(e, V // 2)
        x = oracle.eo_to_coords(idx, parity, dims)
        assert tuple(x) == tuple(coords[e])
        assert oracle.coords_to_eo(x, dims) == e
        for mu in range(4):
            for d in (-1, 1):
                y = list(x)
                y[mu] += d
                assert oracle.neighbor_eo(idx, parity, dims, mu, d) == eo_at(eo_of, y, dims)
            for nu in range(4):
                if nu == mu:
                    continue
                y = list(x)
                y[mu] += 1
                y[nu] -= 1
                assert oracle.neighbor_eo(idx, parity, dims, mu, 1, nu, -1) == eo_at(eo_of, y, dims)


@pytest.fixture(scope="module")
def geom_check():
    exe = os.path.join(ROOT, "_build", "geom_check")
    src = os.path.join(ROOT, "tests", "host", "geom_check.cpp")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT}/sun_amd/csrc", f"-I{ROOT}/include", src, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("dims", DIMS)
def test_device_geometry_matches_brute_force(geom_check, dims):
    out = subprocess.run([geom_check, *map(str, dims)], capture_output=True, text=True, check=True).stdout
    rows = np.array([list(map(int, l.split())) for l in out.strip().splitlines()], dtype=np.int64)
    coords, eo_of = brute(dims)
    V = int(np.prod(dims))
    assert rows.shape == (V, 6 + 8 + 12)
    neigh = rows[:, 6:]
    assert neigh.min() >= 0 and neigh.max() < V
    for r in rows:
        parity, idx = int(r[0]), int(r[1])
        e = parity * (V // 2) + idx
        x = tuple(int(v) for v in r[2:6])
        assert x == tuple(coords[e])
        k = 6
        for mu in range(4):
            for d in (-1, 1):
                y = list(x)
                y[mu] += d
                assert r[k] == eo_at(eo_of, y, dims)
                k += 1
        for mu in range(4):
            for nu in range(4):
                if mu == nu:
                    continue
                y = list(x)
                y[nu] -= 1
                y[mu] += 1
                assert r[k] == eo_at(eo_of, y, dims)
                k += 1


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 8, 10)])
def test_sub_sweep_is_race_free(oracle, dims):
    """Read set of a (parity, mu) launch never contains another site's written link
    (the argument the reference relies on at monte.cu:64-80)."""
    V = int(np.prod(dims))
    half = V // 2
    for parity in (0, 1):
        for mu in range(4):
            written = {mu * V + parity * half + i for i in range(half)}
            for i in range(half):
                own = mu * V + parity * half + i
                reads = set()
                pmu = oracle.neighbor_eo(i, parity, dims, mu, 1)
                for nu in range(4):
                    if nu == mu:
                        continue
                    reads.add(nu * V + parity * half + i)
                    reads.add(mu * V + oracle.neighbor_eo(i, parity, dims, nu, 1))
                    reads.add(nu * V + pmu)
                    mnu = oracle.neighbor_eo(i, parity, dims, nu, -1)
                    reads.add(nu * V + mnu)
                    reads.add(mu * V + mnu)
                    reads.add(nu * V + oracle.neighbor_eo(i, parity, dims, mu, 1, nu, -1))
                assert not ((reads & written) - {own})
