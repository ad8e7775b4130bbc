"""Pinning of the oracle's link monitor (linktrsum.cu:130-150, linkdetsum.cu:84-101).

* the mean trace / N, mean determinant and the two unitarity deviations equal numpy's
  trace, linalg.det (LU with pivoting, an independent algorithm) and the Frobenius
  entries of U U^+ - 1 on random complex (non-unitary) fields;
* closed forms: the identity field gives (1, 0, 1, 0, 0, 0) exactly; diagonal phase
  links diag(e^{i t_k}) give trace sum_k e^{i t_k} / N and det e^{i sum t_k};
* on Haar-random SU(N) fields det = 1 and U U^+ = 1 to rounding, and the mean trace
  is zero within its statistical error (E Tr U = 0, E |Tr U|^2 = 1 for Haar SU(N)).
A planted mistake in the normalisation (4V vs 4NV), a dropped imaginary part or a
transposed product in U U^+ (harmless only for normal matrices) fails one of these.
"""
import numpy as np
import pytest

NCS = [2, 3, 4]


def numpy_monitor(U):
    nc = U.shape[-1]
    L = U.reshape(-1, nc, nc)
    tr = np.trace(L, axis1=1, axis2=2).mean() / nc
    dt = np.linalg.det(L)
    G = L @ np.conj(np.transpose(L, (0, 2, 1))) - np.eye(nc)
    return np.array([tr.real, tr.imag, dt.mean().real, dt.mean().imag, np.abs(G).max(), np.abs(dt - 1).max()])


@pytest.mark.parametrize("nc", NCS)
def test_random_fields_match_numpy(oracle, nc):
    rng = np.random.default_rng(40 + nc)
    U = (rng.standard_normal((4, 48, nc, nc)) + 1j * rng.standard_normal((4, 48, nc, nc))) / np.sqrt(2)
    got, ref = oracle.link_monitor(U), numpy_monitor(U)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13), (got, ref)


@pytest.mark.parametrize("nc", NCS)
def test_non_normal_matrix_catches_transposed_product(oracle, nc):
    # an upper-triangular unipotent matrix: U U^+ != U^+ U, det 1
    A = np.eye(nc, dtype=complex)
    A[0, nc - 1] = 0.5 + 0.25j
    U = np.broadcast_to(A, (4, 2, nc, nc)).copy()
    got = oracle.link_monitor(U)
    G = A @ A.conj().T - np.eye(nc)
    assert got[4] == pytest.approx(np.abs(G).max(), abs=1e-15)
    assert got[5] == pytest.approx(0.0, abs=1e-15)
    assert got[2] == pytest.approx(1.0, abs=1e-15)


@pytest.mark.parametrize("nc", NCS)
def test_identity_exact(oracle, nc):
    U = oracle.cold_start(nc, (2, 2, 2, 4))
    assert list(oracle.link_monitor(U)) == [1.0, 0.0, 1.0, 0.0, 0.0, 0.0]


@pytest.mark.parametrize("nc", NCS)
def test_diagonal_phases_closed_form(oracle, nc):
    rng = np.random.default_rng(7)
    th = rng.uniform(-np.pi, np.pi, size=(4, 16, nc))
    U = np.zeros((4, 16, nc, nc), dtype=complex)
    for k in range(nc):
        U[..., k, k] = np.exp(1j * th[..., k])
    got = oracle.link_monitor(U)
    tr = np.exp(1j * th).sum(axis=-1).mean() / nc
    dt = np.exp(1j * th.sum(axis=-1))
    assert got[0] == pytest.approx(tr.real, abs=1e-15) and got[1] == pytest.approx(tr.imag, abs=1e-15)
    assert got[2] == pytest.approx(dt.mean().real, abs=1e-14) and got[3] == pytest.approx(dt.mean().imag, abs=1e-14)
    assert got[4] < 1e-15
    assert got[5] == pytest.approx(np.abs(dt - 1).max(), abs=1e-14)


@pytest.mark.parametrize("nc", NCS)
def test_haar_field(oracle, nc):
    dims = (4, 4, 4, 8)
    U = oracle.hot_start(nc, dims, 11)
    got = oracle.link_monitor(U)
    assert got[4] < 1e-13 and got[5] < 1e-13
    assert got[2] == pytest.approx(1.0, abs=1e-14) and abs(got[3]) < 1e-14
    n = 4 * 512
    sigma = np.sqrt(1.0 / n) / nc  # bounds the std of Re and Im of the mean of Tr U / N
    assert abs(got[0]) < 5 * sigma and abs(got[1]) < 5 * sigma


def test_nan_propagates(oracle):
    U = oracle.cold_start(3, (2, 2, 2, 4))
    U[2, 5, 1, 2] = np.nan
    got = oracle.link_monitor(U)
    # an off-diagonal entry: the trace stays finite, everything else is NaN
    assert np.isfinite(got[:2]).all() and np.isnan(got[2:]).all()
