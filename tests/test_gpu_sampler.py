"""The device's SU(2) subgroup sampler (Kennedy-Pendleton for alpha > 2, Creutz for
alpha <= 2, device_PHB_OVR.h:162-223) against its exact law, on the GPU.

On a cold SU(2) lattice every staple of a link is the identity, so F = 6 * 1, the
projected vector is r = (12, 0, 0, 0) and one heatbath launch replaces each of its
V/2 links by a draw a of the sampler itself at alpha = (beta/2) * 12 = 6 beta:
a0 = Re U'_00 has density proportional to sqrt(1 - a0^2) exp(alpha a0) on [-1, 1]
(mean I2(alpha)/I1(alpha)) and (a1, a2, a3) is isotropic with |a|^2 = 1 - a0^2.
Every site draws from its own Philox stream, so the 32768 links of a 16^4 parity are
independent samples.  FP32 (native log/cos/sin/sqrt) and FP64, both branches.
"""
import numpy as np
import pytest
from scipy import special, stats

import sun_amd

pytestmark = pytest.mark.gpu

DIMS = (16, 16, 16, 16)


def law_cdf(al):
    """CDF of sqrt(1-t^2) exp(al t) on [-1, 1], tabulated on a fine grid."""
    t = np.linspace(-1.0, 1.0, 200001)
    w = np.sqrt(np.clip(1.0 - t * t, 0.0, None)) * np.exp(al * (t - 1.0))
    c = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(t))])
    c /= c[-1]
    return lambda x: np.interp(x, t, c)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("alpha", [0.6, 1.8, 3.0, 9.0, 24.0])
def test_device_sampler_exact_law(ctx, prec, alpha):
    beta = alpha / 6.0
    with sun_amd.Lattice(ctx, 2, DIMS, beta, 18, prec) as lat:
        lat.cold_start()
        lat.heatbath_sub(seed=1234, sweep=int(alpha * 10), parity=0, mu=0)
        U = lat.to_host()
    V = int(np.prod(DIMS))
    new = U[0, : V // 2]                      # (parity 0, mu 0): the updated links
    assert np.all(U[0, V // 2:] == np.eye(2)) and np.all(U[1:] == np.eye(2))
    a0 = new[:, 0, 0].real
    a = np.stack([new[:, 0, 1].imag, new[:, 0, 1].real, new[:, 0, 0].imag], axis=1)  # (a1, a2, a3)
    # the draws are SU(2) elements [[a0 + i a3, a2 + i a1], [-a2 + i a1, a0 - i a3]]
    tol = 2e-6 if prec == 4 else 1e-13
    assert np.abs(new[:, 1, 1] - np.conj(new[:, 0, 0])).max() <= tol
    assert np.abs(new[:, 1, 0] + np.conj(new[:, 0, 1])).max() <= tol
    assert np.abs(a0 ** 2 + (a ** 2).sum(axis=1) - 1.0).max() <= 10 * tol
    n = len(a0)
    mean = special.iv(2, alpha) / special.iv(1, alpha)
    assert abs(a0.mean() - mean) < 5 * a0.std() / np.sqrt(n), (a0.mean(), mean)
    assert stats.kstest(a0, law_cdf(alpha)).pvalue > 1e-4
    # direction of (a1, a2, a3): isotropic -- zero mean, equal second moments
    s2 = (a ** 2).sum(axis=1)
    for i in range(3):
        assert abs(a[:, i].mean()) < 5 * np.sqrt(s2.mean() / 3 / n)
        assert abs((a[:, i] ** 2).mean() - s2.mean() / 3) < 5 * np.sqrt(np.var(a[:, i] ** 2) / n) + 1e-12
