"""Oracle RNG and SU(2) sampler, CPU only.

Philox4x32-10 is pinned by the Random123 known-answer vectors
(tests/golden/philox4x32_10_kat.json).  The KP/Creutz sampler restating
device_PHB_OVR.h:162-223 is pinned by the law it must produce: a0 has density
proportional to sqrt(1-a0^2) exp(al*a0) on [-1,1] (mean I2(al)/I1(al)), and
(a1,a2,a3) is uniform on the sphere of radius sqrt(1-a0^2).
"""
import json
import os

import numpy as np
import pytest
from scipy import integrate, special, stats

HERE = os.path.dirname(os.path.abspath(__file__))


def test_philox_known_answers(oracle):
    kat = json.load(open(os.path.join(HERE, "golden", "philox4x32_10_kat.json")))
    for v in kat["vectors"]:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        out = tuple(int(x, 16) for x in v["out"])
        assert oracle.philox(ctr, key) == out


def test_u01_range_and_edges(oracle):
    assert oracle.u01(0) == 2.0 ** -32
    assert oracle.u01(1) == 2.0 * 2.0 ** -32
    assert oracle.u01(0xFFFFFFFF) == 1.0
    assert oracle.u01(0x80000000) == 0.5 + 2.0 ** -32


def test_stream_layout(oracle):
    """draw k of stream (seed,c0,c2,c3) = word k%4 of Philox((c0, k//4, c2, c3), seed)"""
    seed = 0x0123456789ABCDEF
    u = oracle.stream_uniforms(seed, 17, 5, 0x10000003, 10)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for k in range(10):
        w = oracle.philox([17, k // 4, 5, 0x10000003], key)
        assert u[k] == oracle.u01(w[k % 4])


def test_stream_uniformity(oracle):
    u = oracle.stream_uniforms(99, 3, 0, 0, 200000)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-4


def a0_cdf(al):
    w = lambda t: np.sqrt(max(0.0, 1 - t * t)) * np.exp(al * (t - 1.0))
    Z = integrate.quad(w, -1, 1, limit=200)[0]
    return lambda x: integrate.quad(w, -1, x, limit=200)[0] / Z


@pytest.mark.parametrize("al", [0.3, 1.5, 2.0, 2.5, 6.0, 15.0])
def test_sampler_a0_law(oracle, al):
    n = 40000
    a0 = oracle.sample_a0(al, 7 + int(al * 10), n)
    assert np.all(np.abs(a0) <= 1.0)
    mean = special.iv(2, al) / special.iv(1, al)
    assert abs(a0.mean() - mean) < 5 * a0.std() / np.sqrt(n)
    cdf = a0_cdf(al)
    xs = np.sort(a0)[:: n // 400]
    emp = (np.arange(len(xs)) * (n // 400) + 1) / n
    assert np.max(np.abs(np.array([cdf(x) for x in xs]) - emp)) < 0.015


def test_sampler_sphere_and_consumption(oracle):
    # explicit uniforms: first KP attempt accepted -> 4 + 2 uniforms
    u = [0.5, 0.7, 0.1, 0.01, 0.25, 0.125]
    a, used = oracle.generate_su2(10.0, u)
    assert used == 6
    assert abs(np.sum(a * a) - 1.0) < 1e-15
    xr1, xr2 = np.log(0.5 + 1e-10), np.log(0.7 + 1e-10)
    c = np.cos(2 * np.pi * 0.1)
    d = -(xr2 + xr1 * c * c) / 10.0
    assert a[0] == pytest.approx(1 - d, abs=1e-15)
    r = np.sqrt(abs(1 - a[0] ** 2))
    assert a[3] == pytest.approx((2 * 0.25 - 1) * r, abs=1e-15)
    # rejection with al <= 2 goes to the Creutz branch: 4 + 2 per try + 2
    u = [0.5, 0.5, 0.0, 0.999, 0.9, 0.1, 0.3, 0.3]
    a, used = oracle.generate_su2(1.0, u)
    assert used == 8
    assert abs(np.sum(a * a) - 1.0) < 1e-15
    a0 = 1.0 + np.log(np.exp(-2.0) + (1 - np.exp(-2.0)) * 0.9) / 1.0
    assert a[0] == pytest.approx(a0, abs=1e-15)


def test_sampler_direction_isotropic(oracle):
    """(a1,a2,a3)/|.| uniform on S^2: each component has mean 0 and variance 1/3."""
    vec = []
    for k in range(10000):
        # one independent stream per sample (the sampler may retry and draw more)
        a, _ = oracle.generate_su2(8.0, oracle.stream_uniforms(5, k, 0, 0x30000000, 96))
        v = a[1:] / np.linalg.norm(a[1:])
        vec.append(v)
    vec = np.array(vec)
    assert np.all(np.abs(vec.mean(axis=0)) < 0.03)
    assert np.all(np.abs((vec ** 2).mean(axis=0) - 1 / 3) < 0.02)
