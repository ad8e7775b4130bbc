"""Oracle colour algebra, codecs, reunitarisation and updates, CPU only.

Each check is chosen so that a plausible transcription mistake (a dropped term, a
wrong sign or index, a transposed operand) fails it:
  * matrix product and determinant against numpy;
  * 12-/8-real codecs (reconstruct_12p_8p.h:17-82) round-trip Haar SU(3);
  * reunitarisation (reunitlink.h:16-85) is idempotent on SU(N) and restores
    unitarity and det = 1 of perturbed links;
  * the Haar hot start has the moments of Haar measure: E[Tr U] = 0,
    E[|Tr U|^2] = 1, E[|Tr U|^4] = 2;
  * the staple (monte.cu:63-80) equals an independent lexicographic numpy staple;
  * overrelaxation (device_PHB_OVR.h:493-608) keeps Re Tr(U F) of every link;
  * the SU(2) heatbath (generic path, device_PHB_OVR.h:426-441) makes U'F = k a
    with a an SU(2) element whose a0 has the exact single-link law at al = beta*k.
"""
import numpy as np
import pytest
from scipy import special, stats

NCS = [2, 3, 4]


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def unitarity_err(U):
    n = U.shape[-1]
    I = np.einsum("...ij,...kj->...ik", U, U.conj())
    return np.abs(I - np.eye(n)).max()


@pytest.mark.parametrize("nc", NCS)
def test_matmul_and_det(oracle, nc):
    rng = np.random.default_rng(nc)
    for _ in range(20):
        A, B = rand_c(rng, nc, nc), rand_c(rng, nc, nc)
        assert np.abs(oracle.matmul(A, B) - A @ B).max() < 1e-13
        assert abs(oracle.det(A) - np.linalg.det(A)) < 1e-12 * max(1.0, abs(np.linalg.det(A)))


def test_codecs_roundtrip(oracle):
    H = oracle.hot_start(3, (4, 4, 4, 4), 3)
    links = H.reshape(-1, 3, 3)
    for L in links[:500]:
        assert np.abs(oracle.decode12(oracle.encode12(L)) - L).max() < 1e-14
        p8 = oracle.encode8(L)
        assert np.abs(oracle.decode8(p8) - L).max() < 5e-14
        assert p8[6] == pytest.approx(np.angle(L[0, 0]), abs=1e-15)
        assert p8[7] == pytest.approx(np.angle(L[2, 0]), abs=1e-15)


def test_decode12_is_cross_product(oracle):
    rng = np.random.default_rng(1)
    p = rng.standard_normal(12)
    M = oracle.decode12(p)
    r0 = p[0:6:2] + 1j * p[1:6:2]
    r1 = p[6:12:2] + 1j * p[7:12:2]
    assert np.allclose(M[0], r0) and np.allclose(M[1], r1)
    assert np.abs(M[2] - np.conj(np.cross(r0, r1))).max() < 1e-14


@pytest.mark.parametrize("nc", NCS)
def test_reunit(oracle, nc):
    H = oracle.hot_start(nc, (2, 2, 2, 2), 4).reshape(-1, nc, nc)
    rng = np.random.default_rng(2)
    for L in H[:40]:
        R = oracle.reunit_link(L)
        assert np.abs(R - L).max() < 5e-14                      # idempotent on SU(N)
        P = L + 0.05 * rand_c(rng, nc, nc)
        R = oracle.reunit_link(P)
        assert unitarity_err(R) < 1e-14
        assert abs(np.linalg.det(R) - 1.0) < 1e-14
        assert np.abs(R - L).max() < 0.5                      # stays near the input


def test_reunit_su3_rows(oracle):
    rng = np.random.default_rng(3)
    P = rand_c(rng, 3, 3)
    R = oracle.reunit_link(P)
    r0 = P[0] / np.linalg.norm(P[0])
    assert np.abs(R[0] - r0).max() < 1e-15
    r1 = P[1] - np.vdot(r0, P[1]) * r0
    r1 /= np.linalg.norm(r1)
    assert np.abs(R[1] - r1).max() < 1e-14
    assert np.abs(R[2] - np.conj(np.cross(R[0], R[1]))).max() < 1e-15


@pytest.mark.parametrize("nc", NCS)
def test_hot_start_haar_moments(oracle, nc):
    """Haar moments of the trace: E Tr U = 0, E|Tr U|^2 = 1, E|Tr U|^4 = 2 (N >= 2),
    each within 5 sigma; Var|Tr U|^4 = E|Tr U|^8 - 4 <= 20 for SU(2..4)."""
    H = oracle.hot_start(nc, (12, 12, 12, 12), 10 + nc).reshape(-1, nc, nc)
    # worst case over 83k Gram-Schmidt orthonormalisations of Gaussian matrices
    assert unitarity_err(H) < 1e-12
    assert np.abs(np.linalg.det(H) - 1).max() < 1e-12
    tr = np.trace(H, axis1=1, axis2=2)
    n = len(tr)
    assert abs(tr.mean()) < 5 / np.sqrt(n)
    m2 = np.mean(np.abs(tr) ** 2)
    m4 = np.mean(np.abs(tr) ** 4)
    assert abs(m2 - 1.0) < 5 * np.sqrt(2.0 / n) + 0.01
    assert abs(m4 - 2.0) < 5 * np.sqrt(20.0 / n)


def lex_field(oracle, U, dims):
    """canonical EO field -> array [mu, x3, x2, x1, x0, N, N]"""
    V = int(np.prod(dims))
    nc = U.shape[-1]
    out = np.zeros((4, dims[3], dims[2], dims[1], dims[0], nc, nc), dtype=complex)
    for x3 in range(dims[3]):
        for x2 in range(dims[2]):
            for x1 in range(dims[1]):
                for x0 in range(dims[0]):
                    e = oracle.coords_to_eo((x0, x1, x2, x3), dims)
                    out[:, x3, x2, x1, x0] = U[:, e]
    return out


def shift(A, mu, d):
    """A(x + d*e_mu) for arrays indexed [x3, x2, x1, x0, ...]"""
    axis = 3 - mu
    return np.roll(A, -d, axis=axis)


def numpy_staple(L, mu):
    """sum_{nu != mu} U_nu(x)U_mu(x+nu)U_nu(x+mu)^+ + U_nu(x-nu)^+U_mu(x-nu)U_nu(x-nu+mu)"""
    dag = lambda M: np.conj(np.swapaxes(M, -1, -2))
    S = np.zeros_like(L[0])
    for nu in range(4):
        if nu == mu:
            continue
        S += L[nu] @ shift(L[mu], nu, 1) @ dag(shift(L[nu], mu, 1))
        Ld = shift(L[nu], nu, -1)
        S += dag(Ld) @ shift(L[mu], nu, -1) @ shift(shift(L[nu], nu, -1), mu, 1)
    return S


@pytest.mark.parametrize("nc", NCS)
def test_staple_brute_force(oracle, nc):
    dims = (4, 2, 4, 6)
    U = oracle.hot_start(nc, dims, 6)
    L = lex_field(oracle, U, dims)
    V = int(np.prod(dims))
    for mu in range(4):
        S = numpy_staple(L, mu)
        for parity in (0, 1):
            for idx in range(0, V // 2, 3):
                x = oracle.eo_to_coords(idx, parity, dims)
                ref = S[x[3], x[2], x[1], x[0]]
                got = oracle.staple(U, dims, parity, mu, idx)
                assert np.abs(got - ref).max() < 1e-12


@pytest.mark.parametrize("nc", NCS)
def test_overrelax_keeps_local_action(oracle, nc):
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(nc, dims, 9)
    rng = np.random.default_rng(nc)
    for k in range(60):
        mu, parity, idx = k % 4, (k // 4) % 2, int(rng.integers(0, 128))
        S = oracle.staple(U, dims, parity, mu, idx)
        F = np.conj(S.T)
        link = U[mu, parity * 128 + idx]
        new = oracle.overrelax_link(link, F)
        a_old = np.trace(link @ F).real
        a_new = np.trace(new @ F).real
        assert abs(a_new - a_old) < 1e-12 * max(1, abs(a_old))
        assert unitarity_err(new) < 1e-13
        assert np.abs(new - link).max() > 1e-3          # it does move the link


def test_su2_heatbath_exact_law(oracle):
    """SU(2): U' F = k a with a in SU(2), a0 ~ sqrt(1-a0^2) exp(beta k a0)."""
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(2, dims, 12)
    S = oracle.staple(U, dims, 0, 1, 5)
    F = np.conj(S.T)
    link = U[1, 5]
    beta = 0.7
    M = link @ F
    r = np.array([M[0, 0].real + M[1, 1].real, M[0, 1].imag + M[1, 0].imag,
                  M[0, 1].real - M[1, 0].real, M[0, 0].imag - M[1, 1].imag])
    k = np.linalg.norm(r)
    al = beta / 2 * k
    a0 = []
    for i in range(8000):
        new = oracle.heatbath_link(link, F, beta, 77, c0=i)
        W = new @ F / (k / 2)
        assert unitarity_err(W) < 1e-12 and abs(np.linalg.det(W) - 1) < 1e-12
        a0.append(W[0, 0].real)
    a0 = np.array(a0)
    mean = special.iv(2, al) / special.iv(1, al)
    assert abs(a0.mean() - mean) < 5 * a0.std() / np.sqrt(len(a0))
    assert stats.kstest(a0, lambda x: law_cdf(al, x)).pvalue > 1e-4


def law_cdf(al, xs):
    from scipy import integrate
    w = lambda t: np.sqrt(max(0.0, 1 - t * t)) * np.exp(al * (t - 1.0))
    Z = integrate.quad(w, -1, 1)[0]
    xs = np.atleast_1d(xs)
    return np.array([integrate.quad(w, -1, x)[0] / Z for x in xs])


@pytest.mark.parametrize("nc", NCS)
def test_heatbath_link_stays_in_sun(oracle, nc):
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(nc, dims, 13)
    for k in range(20):
        S = oracle.staple(U, dims, k % 2, k % 4, 7 * k % 128)
        new = oracle.heatbath_link(U[k % 4, 7 * k % 128], np.conj(S.T), 6.0, 5, c0=k)
        assert unitarity_err(new) < 1e-13
        assert abs(np.linalg.det(new) - 1) < 1e-13


def test_portable_transcendentals_against_glibc(tmp_path):
    """sun_libm.h -- the FP64 log, exp, sin/cos and atan2 that the device sampler,
    Haar start and 8-real codec share with the oracle -- within 2 ulp of glibc's log
    and exp (sampler arguments and the whole exponent range), within 3 ulp of its
    atan2 (all quadrants; the pi/4 + atan((t-1)/(t+1)) step cancels just above
    tan(pi/8)),
    within 2 x 2^-53 of long-double sin/cos of 2 pi u and of phases in [-pi, pi],
    with exact values at the quadrant points, log(2^k) = k ln 2 and C's conventions
    for zeros, NaN and overflow"""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "libm_check")
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", f"-I{root}/sun_amd/csrc",
                    os.path.join(root, "tests", "host", "libm_check.c"), "-o", exe, "-lm"], check=True)
    mlog, mexp, msin, mcos, matan, nfail = subprocess.run([exe], capture_output=True, text=True,
                                                          check=True).stdout.split()
    assert float(mlog) <= 2 and float(mexp) <= 2 and float(matan) <= 3, (mlog, mexp, matan)
    assert float(msin) <= 2 and float(mcos) <= 2, (msin, mcos)
    assert int(nfail) == 0
