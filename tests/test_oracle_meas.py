"""Pinning of the oracle's measurement chain: SU(N) projection (projectlink.h),
APE smearing (ape.cu), multihit (multihit.cu) and Wilson loops (wilsonloop.cu).

* projection: F = c W (W in SU(3), c > 0) projects to W; for random F the result is
  in SU(3) and a local maximum of Re Tr(U F^+) (no small rotation increases it);
  for SU(2) one hit gives the closed-form maximiser, the normalised quaternion part;
* APE: an independent numpy implementation for SU(2) (lexicographic field, np.roll
  staples, closed-form SU(2) projection) agrees link by link; SU(2) smearing is
  exactly gauge covariant; the identity field is a fixed point; smearing raises the
  spatial plaquette;
* multihit: nhit = 0 is the identity; nhit = 1 is (U + HB(U)) / 2 with the pinned
  heatbath link update on the same stream; for SU(2) the mean over hits approaches
  the exact conditional mean I2(a)/I1(a) V^+ (a = beta/2 |F|), within its error;
* Wilson loops: the identity field gives 1; W(r, 0) = W(0, T) = 1 on SU(N) fields;
  W(1, 1) is the temporal plaquette (pinned separately); W is gauge invariant; an
  abelian field with temporal phases exp(i phi_k x0) gives the Stokes closed form
  (mean_k exp(i phi_k r T) + 2) / 3.
"""
import numpy as np
import pytest
from scipy import special

NCS = [2, 3, 4]


def haar(rng, n, shape=()):
    Z = (rng.standard_normal(shape + (n, n)) + 1j * rng.standard_normal(shape + (n, n))) / np.sqrt(2)
    Q, Rm = np.linalg.qr(Z)
    d = np.diagonal(Rm, axis1=-2, axis2=-1)
    Q = Q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(Q)
    return Q / (det ** (1.0 / n))[..., None, None]


def retr(U, F):
    return np.real(np.trace(U @ F.conj().T))


def test_su3_project_fixed_point(oracle):
    rng = np.random.default_rng(1)
    for _ in range(10):
        W = haar(rng, 3)
        got = oracle.su3_project(2.7 * W, 200)
        assert np.abs(got - W).max() < 1e-10


def test_su3_project_local_maximum(oracle):
    rng = np.random.default_rng(2)
    for _ in range(10):
        F = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        U = oracle.su3_project(F, 300)
        assert np.abs(U @ U.conj().T - np.eye(3)).max() < 1e-13
        assert abs(np.linalg.det(U) - 1) < 1e-13
        f0 = retr(U, F)
        for _ in range(20):
            H = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
            H = (H + H.conj().T) / 2
            H -= np.trace(H) / 3 * np.eye(3)
            w, v = np.linalg.eigh(H)
            E = (v * np.exp(1e-3j * w)) @ v.conj().T
            assert retr(E @ U, F) <= f0 + 1e-12
        for _ in range(50):
            assert retr(haar(rng, 3), F) <= f0 + 1e-12


def quat_project_su2(F):
    """maximiser of Re Tr(U F^+) over SU(2): the normalised quaternion part of F"""
    a0 = (F[0, 0].real + F[1, 1].real) / 2
    a3 = (F[0, 0].imag - F[1, 1].imag) / 2
    a2 = (F[0, 1].real - F[1, 0].real) / 2
    a1 = (F[0, 1].imag + F[1, 0].imag) / 2
    n = np.sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3)
    a0, a1, a2, a3 = a0 / n, a1 / n, a2 / n, a3 / n
    return np.array([[a0 + 1j * a3, a2 + 1j * a1], [-a2 + 1j * a1, a0 - 1j * a3]])


def test_project_link_su2_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        F = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        w = haar(rng, 2)
        assert np.abs(oracle.project_link(w, F, 1) - quat_project_su2(F)).max() < 1e-14


@pytest.mark.parametrize("nc", [2, 4])
def test_project_link_tolerance_stops_early(oracle, nc):
    """with tol > 0 the hits stop after the first one whose relative change of
    Re Tr(w F) / N (projectlink.h:22,57: realtraceUVdagger of w and the daggered F)
    is at most tol"""
    rng = np.random.default_rng(4)
    for tol in (1e-2, 1e-4, 1e-8):
        F = rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc))
        w = haar(rng, nc)
        prev, m_prev = w, np.real(np.trace(w @ F)) / nc
        for k in range(1, 60):
            wk = oracle.project_link(w, F, k, 0.0)
            m = np.real(np.trace(wk @ F)) / nc
            conv = (m - m_prev) / m_prev
            prev, m_prev = wk, m
            if conv <= tol:
                break
        assert np.array_equal(oracle.project_link(w, F, 60, tol), prev), (tol, k)


# ---- APE: an independent numpy restatement for SU(2) on a lexicographic field ----
def to_lex(U, dims):
    V = int(np.prod(dims))
    X0, X1, X2, X3 = dims
    L = np.empty((4, X3, X2, X1, X0, U.shape[-1], U.shape[-1]), dtype=complex)
    for lex in range(V):
        x0, x1, x2, x3 = lex % X0, (lex // X0) % X1, (lex // (X0 * X1)) % X2, lex // (X0 * X1 * X2)
        eo = lex // 2 + ((x0 + x1 + x2 + x3) & 1) * (V // 2)
        L[:, x3, x2, x1, x0] = U[:, eo]
    return L


def from_lex(L, dims):
    V = int(np.prod(dims))
    X0, X1, X2, X3 = dims
    U = np.empty((4, V) + L.shape[-2:], dtype=complex)
    for lex in range(V):
        x0, x1, x2, x3 = lex % X0, (lex // X0) % X1, (lex // (X0 * X1)) % X2, lex // (X0 * X1 * X2)
        eo = lex // 2 + ((x0 + x1 + x2 + x3) & 1) * (V // 2)
        U[:, eo] = L[:, x3, x2, x1, x0]
    return U


AX = {0: 3, 1: 2, 2: 1, 3: 0}  # direction -> array axis of [x3, x2, x1, x0]


def shift(A, mu, d):
    """A(x + d e_mu)"""
    return np.roll(A, -d, axis=AX[mu])


def dag(A):
    return np.conj(np.swapaxes(A, -1, -2))


def ape_numpy_su2(L, w, steps):
    for _ in range(steps):
        out = L.copy()
        for mu in range(3):
            S = np.zeros_like(L[mu])
            for nu in range(3):
                if nu == mu:
                    continue
                S += L[nu] @ shift(L[mu], nu, 1) @ dag(shift(L[nu], mu, 1))
                Ld = shift(L[nu], nu, -1)
                S += dag(Ld) @ shift(L[mu], nu, -1) @ shift(shift(L[nu], nu, -1), mu, 1)
            M = (L[mu] + w * S) / (1 + 6 * w)
            P = np.empty_like(M)
            for idx in np.ndindex(M.shape[:-2]):
                P[idx] = quat_project_su2(M[idx])
            out[mu] = P
        L = out
    return L


def test_ape_su2_matches_numpy(oracle):
    dims = (4, 4, 4, 6)
    U = oracle.hot_start(2, dims, 5)
    ref = from_lex(ape_numpy_su2(to_lex(U, dims), 0.3, 2), dims)
    got = oracle.ape_smear(U.copy(), dims, 0.3, 2, 1)
    assert np.abs(got - ref).max() < 1e-13


def gauge_transform(U, dims, G):
    """U_mu(x) -> G(x) U_mu(x) G(x + mu)^+ on the even/odd field"""
    L = to_lex(U, dims)
    Gl = to_lex(np.broadcast_to(G, (4,) + G.shape).copy(), dims)[0]
    out = np.empty_like(L)
    for mu in range(4):
        out[mu] = Gl @ L[mu] @ dag(shift(Gl, mu, 1))
    return from_lex(out, dims)


def test_ape_su2_gauge_covariant(oracle):
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(6)
    U = oracle.hot_start(2, dims, 7)
    G = haar(rng, 2, (256,))
    a = oracle.ape_smear(gauge_transform(U, dims, G), dims, 0.2, 3, 1)
    b = gauge_transform(oracle.ape_smear(U.copy(), dims, 0.2, 3, 1), dims, G)
    assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("nc", NCS)
def test_ape_identity_fixed_and_smooths(oracle, nc):
    dims = (4, 4, 4, 4)
    U = oracle.cold_start(nc, dims)
    assert np.abs(oracle.ape_smear(U.copy(), dims, 0.2, 3, 10) - U).max() < 1e-14
    H = oracle.hot_start(nc, dims, 8)
    for it in range(4):  # a rough but not random field
        oracle.heatbath_sweep(H, dims, {2: 2.3, 3: 6.0, 4: 11.0}[nc], 9, it)
    p0 = oracle.plaquette(H, dims)
    S = oracle.ape_smear(H.copy(), dims, 0.2, 5, 10, 1e-10)
    p1 = oracle.plaquette(S, dims)
    assert p1[0] > p0[0] + 0.05
    assert np.abs(S[3] - H[3]).max() == 0.0  # time links untouched by spatial smearing
    T = oracle.ape_smear(H.copy(), dims, 0.2, 1, 10, 1e-10, time_dir=True)
    assert np.abs(T[:3] - H[:3]).max() == 0.0 and np.abs(T[3] - H[3]).max() > 0.01


# ---- multihit ----
def test_multihit_zero_and_one_hit(oracle):
    dims, beta, seed = (4, 4, 4, 4), 6.0, 21
    U = oracle.hot_start(3, dims, 3)
    assert np.array_equal(oracle.multihit(U, dims, beta, seed, 0), U)
    M = oracle.multihit(U, dims, beta, seed, 1)
    assert np.array_equal(M[:3], U[:3])
    V, half = 256, 128
    for idd in (0, 77, 130, 255):
        parity, sid = idd // half, idd % half
        S = oracle.staple(U, dims, parity, 3, sid)
        hb = oracle.heatbath_link(U[3, idd].copy(), S.conj().T, beta, seed, c0=sid, c2=0,
                                  c3=0x40000000 | (parity << 4) | 3)
        assert np.abs(M[3, idd] - (U[3, idd] + hb) / 2).max() < 1e-15


def test_multihit_su2_conditional_mean(oracle):
    dims, beta, nhit = (4, 4, 4, 4), 2.3, 800
    U = oracle.hot_start(2, dims, 4)
    M = oracle.multihit(U, dims, beta, 33, nhit)
    dev = []
    for idd in range(256):
        parity, sid = idd // 128, idd % 128
        F = oracle.staple(U, dims, parity, 3, sid).conj().T  # F = staple^+
        k = np.sqrt(abs(np.linalg.det(F)))  # SU(2) staples are k x SU(2)
        Vm = F / k
        a = beta * k  # weight exp(beta/2 Re Tr(U F)) = exp(beta k x0), U = X Vm^+
        mean = special.iv(2, a) / special.iv(1, a) * Vm.conj().T
        # the output also carries the start link with weight 1 / (nhit + 1)
        expect = (U[3, idd] + nhit * mean) / (nhit + 1)
        dev.append((M[3, idd] - expect).ravel())
    dev = np.array(dev)
    # per component the mean over hits has std <= 1 / sqrt(nhit); over 256 links
    assert np.abs(dev.mean(axis=0)).max() < 5 / np.sqrt(nhit * 256)
    assert np.sqrt((np.abs(dev) ** 2).mean()) < 2 / np.sqrt(nhit)


# ---- Wilson loops ----
@pytest.mark.parametrize("nc", NCS)
def test_wilson_identity_and_trivial_extents(oracle, nc):
    dims = (4, 4, 6, 6)
    W = oracle.wilson_loops(oracle.cold_start(nc, dims), dims, 3, 4)
    assert np.abs(W - 1).max() < 1e-15
    H = oracle.hot_start(nc, dims, 12)
    W = oracle.wilson_loops(H, dims, 3, 4)
    assert np.abs(W[:, 0] - 1).max() < 1e-13 and np.abs(W[0, :] - 1).max() < 1e-13
    assert W[1, 1].real == pytest.approx(oracle.plaquette(H, dims)[1], abs=1e-14)


def test_wilson_gauge_invariant(oracle):
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(13)
    U = oracle.hot_start(3, dims, 14)
    for it in range(3):
        oracle.heatbath_sweep(U, dims, 6.0, 15, it)
    G = haar(rng, 3, (384,))
    a = oracle.wilson_loops(U, dims, 3, 4)
    b = oracle.wilson_loops(gauge_transform(U, dims, G), dims, 3, 4)
    assert np.abs(a - b).max() < 1e-13
    assert np.abs(a[2:, 2:]).max() < 0.5  # not trivially 1


def test_wilson_abelian_stokes(oracle):
    dims = (8, 4, 4, 6)
    n = np.array([1, 2, -3])
    phi = 2 * np.pi * n / dims[0]
    U = oracle.cold_start(3, dims)
    V = int(np.prod(dims))
    for s in range(V):
        parity = s // (V // 2)
        x = oracle.eo_to_coords(s % (V // 2), parity, dims)
        U[3, s] = np.diag(np.exp(1j * phi * x[0]))
    W = oracle.wilson_loops(U, dims, 4, 5)
    r = np.arange(5)[:, None]
    T = np.arange(6)[None, :]
    expect = (np.exp(1j * phi[None, None, :] * (r * T)[..., None]).mean(axis=-1) + 2) / 3
    assert np.abs(W - expect).max() < 1e-14


# ---- stout smearing (Morningstar-Peardon exponential) ----
def herm_traceless(rng, scale):
    H = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    H = (H + H.conj().T) / 2
    H -= np.trace(H) / 3 * np.eye(3)
    return H * scale / np.linalg.norm(H)


def test_exp_iq_matches_expm(oracle):
    from scipy.linalg import expm
    rng = np.random.default_rng(30)
    cases = []
    for scale in (1e-9, 1e-4, 0.03, 0.0894, 0.0895, 0.2, 1.0, 3.0, 7.0):  # both sides of c1 = 4e-3
        for _ in range(20):
            cases.append(herm_traceless(rng, scale))
    for a in (1e-6, 0.05, 0.7, 2.0):  # two equal eigenvalues: c0 = +-c0max
        cases.append(np.diag([a, a, -2 * a]).astype(complex))
        cases.append(np.diag([-a, -a, 2 * a]).astype(complex))
    cases.append(np.zeros((3, 3), dtype=complex))
    for Q in cases:
        E = oracle.exp_iq(Q)
        R = expm(1j * Q)
        assert np.abs(E - R).max() < 1e-13 * max(1.0, np.abs(Q).max()), (np.abs(E - R).max(), np.abs(Q).max())
        # random unitary conjugation keeps the eigenvalue structure (c0, c1)
        W = haar(rng, 3)
        assert np.abs(oracle.exp_iq(W @ Q @ W.conj().T) - W @ E @ W.conj().T).max() < 1e-12


def test_stout_gauge_covariant_unitary_and_fixed_point(oracle):
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(31)
    U = oracle.hot_start(3, dims, 32)
    for it in range(3):
        oracle.heatbath_sweep(U, dims, 6.0, 33, it)
    G = haar(rng, 3, (384,))
    a = oracle.stout_smear(gauge_transform(U, dims, G), dims, 0.15, 3)
    b = gauge_transform(oracle.stout_smear(U.copy(), dims, 0.15, 3), dims, G)
    assert np.abs(a - b).max() < 1e-12
    S = oracle.stout_smear(U.copy(), dims, 0.15, 3)
    L = S.reshape(-1, 3, 3)
    assert np.abs(L @ np.conj(np.swapaxes(L, 1, 2)) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(L) - 1).max() < 1e-13
    assert np.array_equal(S[3], U[3])
    assert oracle.plaquette(S, dims)[0] > oracle.plaquette(U, dims)[0] + 0.05
    assert np.abs(oracle.stout_smear(U.copy(), dims, 0.0, 2) - U).max() < 1e-15
    C = oracle.cold_start(3, dims)
    assert np.abs(oracle.stout_smear(C.copy(), dims, 0.2, 2) - C).max() < 1e-15
    T = oracle.stout_smear(U.copy(), dims, 0.15, 1, time_dir=True)
    assert np.array_equal(T[:3], U[:3]) and np.abs(T[3] - U[3]).max() > 1e-3


def test_stout_small_rho_is_the_lie_algebra_step(oracle):
    """for small rho, W = (1 + i rho Q1 + O(rho^2)) U with Q1 the traceless Hermitian
    part of -i S U^+ (MP eq. (2)-(3)): checked on one link against a direct numpy
    evaluation of the first-order term"""
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(3, dims, 34)
    rho = 1e-6
    S_ = oracle.stout_smear(U.copy(), dims, rho, 1)
    for idd in (0, 100, 200):
        for mu in range(3):
            Om = rho * spatial_staple(U, dims, mu, idd) @ U[mu, idd].conj().T
            Q = 0.5j * (Om.conj().T - Om)
            Q -= np.trace(Q) / 3 * np.eye(3)
            W1 = (np.eye(3) + 1j * Q) @ U[mu, idd]
            assert np.abs(S_[mu, idd] - W1).max() < 1e-10


def spatial_staple(U, dims, mu, idd):
    L = to_lex(U, dims)
    V = int(np.prod(dims))
    X0, X1, X2, X3 = dims
    # coordinates of idd
    parity, sid = idd // (V // 2), idd % (V // 2)
    found = None
    for lex in range(V):
        x0, x1, x2, x3 = lex % X0, (lex // X0) % X1, (lex // (X0 * X1)) % X2, lex // (X0 * X1 * X2)
        if lex // 2 + ((x0 + x1 + x2 + x3) & 1) * (V // 2) == idd:
            found = (x3, x2, x1, x0)
            break
    S = np.zeros((3, 3), dtype=complex)
    for nu in range(3):
        if nu == mu:
            continue
        Lnu_x = L[nu][found]
        S += Lnu_x @ shift(L[mu], nu, 1)[found] @ dag(shift(L[nu], mu, 1)[found])
        S += dag(shift(L[nu], nu, -1)[found]) @ shift(L[mu], nu, -1)[found] @ shift(shift(L[nu], nu, -1), mu, 1)[found]
    return S


# ---- HYP smearing (hyp.cu): the oracle's recursion against a level-by-level numpy SU(2) ----
def proj_field_su2(M):
    P = np.empty_like(M)
    for idx in np.ndindex(M.shape[:-2]):
        P[idx] = quat_project_su2(M[idx])
    return P


def hyp_numpy_su2(L, a1, a2, a3, mus):
    """one HYP step of the directions mus on a lexicographic SU(2) field; with one hit
    project_link is the closed-form projection whatever its start"""
    def staple(A_mu, A_nu, mu, nu):  # A_nu(x) A_mu(x+nu) A_nu(x+mu)^+ + down
        up = A_nu @ shift(A_mu, nu, 1) @ dag(shift(A_nu, mu, 1))
        dn = dag(shift(A_nu, nu, -1)) @ shift(A_mu, nu, -1) @ shift(shift(A_nu, nu, -1), mu, 1)
        return up + dn
    bar = {}
    for mu in range(4):
        for eta in range(4):
            if eta != mu:
                bar[mu, eta] = proj_field_su2(0.5 * a3 * staple(L[mu], L[eta], mu, eta) + (1 - a3) * L[mu])
    til = {}
    for mu in range(4):
        for nu in range(4):
            if nu == mu:
                continue
            s = np.zeros_like(L[mu])
            for rho in range(4):
                if rho in (mu, nu):
                    continue
                sig = 6 - mu - nu - rho
                s += staple(bar[mu, sig], bar[rho, sig], mu, rho)
            til[mu, nu] = proj_field_su2(0.25 * a2 * s + (1 - a2) * L[mu])
    out = L.copy()
    for mu in mus:
        s = np.zeros_like(L[mu])
        for nu in range(4):
            if nu != mu:
                s += staple(til[mu, nu], til[nu, mu], mu, nu)
        out[mu] = proj_field_su2(a1 / 6 * s + (1 - a1) * L[mu])
    return out


@pytest.mark.parametrize("time_dir", [False, True])
def test_hyp_su2_matches_level_by_level_numpy(oracle, time_dir):
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(2, dims, 40)
    for it in range(2):
        oracle.heatbath_sweep(U, dims, 2.3, 41, it)
    ref = from_lex(hyp_numpy_su2(to_lex(U, dims), 0.75, 0.6, 0.3, [3] if time_dir else [0, 1, 2]), dims)
    got = oracle.hyp_smear(U.copy(), dims, (0.75, 0.6, 0.3), 1, 1, 0.0, time_dir)
    assert np.abs(got - ref).max() < 1e-13


def test_hyp_su2_gauge_covariant(oracle):
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(42)
    U = oracle.hot_start(2, dims, 43)
    G = haar(rng, 2, (256,))
    a = oracle.hyp_smear(gauge_transform(U, dims, G), dims, (0.75, 0.6, 0.3), 2, 1)
    b = gauge_transform(oracle.hyp_smear(U.copy(), dims, (0.75, 0.6, 0.3), 2, 1), dims, G)
    assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("nc", NCS)
def test_hyp_fixed_points_and_smoothing(oracle, nc):
    dims = (4, 4, 4, 4)
    C = oracle.cold_start(nc, dims)
    assert np.abs(oracle.hyp_smear(C.copy(), dims, (0.75, 0.6, 0.3), 1, 10, 1e-10) - C).max() < 1e-14
    U = oracle.hot_start(nc, dims, 44)
    for it in range(3):
        oracle.heatbath_sweep(U, dims, {2: 2.3, 3: 6.0, 4: 11.0}[nc], 45, it)
    assert np.abs(oracle.hyp_smear(U.copy(), dims, (0.0, 0.0, 0.0), 1, 10, 1e-10) - U).max() < 1e-13
    S = oracle.hyp_smear(U.copy(), dims, (0.75, 0.6, 0.3), 1, 10, 1e-10)
    assert np.array_equal(S[3], U[3])
    assert oracle.plaquette(S, dims)[0] > oracle.plaquette(U, dims)[0] + 0.05
    L = S.reshape(-1, nc, nc)
    assert np.abs(L @ np.conj(np.swapaxes(L, 1, 2)) - np.eye(nc)).max() < 1e-12


# ---- spatial multihit (multihitsp.cu): mean of successive overrelaxation steps ----
def test_multihit_space_closed_forms(oracle):
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(3, dims, 50)
    M1 = oracle.multihit_space(U, dims, 1)
    assert np.array_equal(M1[3], U[3])
    for idd in (0, 77, 200):
        for mu in range(3):
            F = spatial_staple(U, dims, mu, idd).conj().T
            # the numpy staple sums in another order: rounding, amplified by the
            # reflection on a Haar field
            assert np.abs(M1[mu, idd] - oracle.overrelax_link(U[mu, idd].copy(), F)).max() < 1e-13
    # SU(2) overrelaxation is a reflection: two steps return the link, so every even
    # nhit gives (OR(U) + U) / 2
    W = oracle.hot_start(2, dims, 51)
    O1 = oracle.multihit_space(W, dims, 1)
    for n in (2, 4):
        Mn = oracle.multihit_space(W, dims, n)
        assert np.abs(Mn[:3] - (O1[:3] + W[:3]) / 2).max() < 1e-14
    with pytest.raises(ValueError):
        oracle.multihit_space(W, dims, 0)


# ---- extended multihit (multihitext.cu): local heatbath sweeps of a 3^3 box ----
def eo_index(x, dims):
    X0, X1, X2, X3 = dims
    lex = x[0] + X0 * (x[1] + X1 * (x[2] + X2 * x[3]))
    return lex // 2 + (sum(x) & 1) * (int(np.prod(dims)) // 2)


def multihit_ext_extended_lattice(o, U, dims, beta, seed, nhit):
    """The reference's own scheme (multihitext.cu:65-108, :122-144, :241-330, :395-429)
    in numpy: a halo-padded copy (padding 2 in every direction) is refreshed from the
    input before each of the 4*4*4*2 sub-lattice passes; in a pass every site of the
    sub-lattice heatbaths the temporal links of its box in place in the copy.  The
    single-link update is the pinned heatbath_link on stream
    (seed, eo(x), hit, 0x50000000 | box index)."""
    nc = U.shape[-1]
    L = to_lex(U, dims)  # [mu, x3, x2, x1, x0]
    lgrid, pad = (4, 4, 4, 2), 2
    blocks = [(dims[d] + lgrid[d] - 1) // lgrid[d] for d in range(4)]
    grid = [blocks[d] * lgrid[d] + 2 * pad for d in range(4)]
    # extended index e holds the input at (e - pad) mod X
    idx = [(np.arange(grid[d]) - pad) % dims[d] for d in range(4)]
    out = np.array(L)
    acc = np.zeros((dims[3], dims[2], dims[1], dims[0], nc, nc), dtype=complex)

    def ext(C, mu, e):
        return C[mu, e[3], e[2], e[1], e[0]]

    for sub in range(lgrid[0] * lgrid[1] * lgrid[2] * lgrid[3]):
        lid = (sub % 4, (sub // 4) % 4, (sub // 16) % 4, sub // 64)
        C = L[:, idx[3]][:, :, idx[2]][:, :, :, idx[1]][:, :, :, :, idx[0]].copy()
        for b3 in range(blocks[3]):
            for b2 in range(blocks[2]):
                for b1 in range(blocks[1]):
                    for b0 in range(blocks[0]):
                        x = [lid[0] + 4 * b0, lid[1] + 4 * b1, lid[2] + 4 * b2, lid[3] + 2 * b3]
                        if any(x[d] >= dims[d] for d in range(4)):
                            continue
                        e = [x[d] + pad for d in range(4)]
                        site = eo_index(x, dims)
                        newU = np.zeros((nc, nc), dtype=complex)
                        for hit in range(nhit):
                            for parity in range(2):
                                for k in (-1, 0, 1):
                                    for j in (-1, 0, 1):
                                        for i in (-1, 0, 1):
                                            if ((i + j + k) & 1) != parity:
                                                continue
                                            y = [e[0] + i, e[1] + j, e[2] + k, e[3]]
                                            S = np.zeros((nc, nc), dtype=complex)
                                            for nu in range(3):
                                                yn = list(y); yn[nu] += 1
                                                yt = list(y); yt[3] += 1
                                                ym = list(y); ym[nu] -= 1
                                                ymt = list(ym); ymt[3] += 1
                                                S += ext(C, nu, y) @ ext(C, 3, yn) @ dag(ext(C, nu, yt))
                                                S += dag(ext(C, nu, ym)) @ ext(C, 3, ym) @ ext(C, nu, ymt)
                                            b = (i + 1) + 3 * (j + 1) + 9 * (k + 1)
                                            C[3, y[3], y[2], y[1], y[0]] = o.heatbath_link(
                                                ext(C, 3, y).copy(), dag(S), beta, seed, c0=site, c2=hit,
                                                c3=0x50000000 | b)
                            newU += ext(C, 3, e)
                        acc[x[3], x[2], x[1], x[0]] = newU / nhit
    out[3] = acc
    return from_lex(out, dims)


@pytest.mark.parametrize("nc,dims", [(2, (4, 4, 4, 2)), (3, (4, 2, 4, 4)), (4, (2, 4, 4, 2))])
def test_multihit_ext_matches_extended_lattice_restatement(oracle, nc, dims):
    U = oracle.hot_start(nc, dims, 60 + nc)
    beta = {2: 2.4, 3: 5.9, 4: 10.8}[nc]
    M = oracle.multihit_ext(U, dims, beta, 77, 2)
    R = multihit_ext_extended_lattice(oracle, U, dims, beta, 77, 2)
    assert np.array_equal(M[:3], U[:3])
    # the numpy staple sums in another order: rounding only, no accept/reject flips
    assert np.abs(M[3] - R[3]).max() < 1e-12


def test_multihit_ext_cold_large_beta_and_unitary_single_hit(oracle):
    dims = (4, 4, 4, 4)
    # cold field, beta huge: every box update stays at the staple's direction (1)
    C = oracle.cold_start(3, dims)
    M = oracle.multihit_ext(C, dims, 1e7, 5, 3)
    assert np.abs(M[3] - np.eye(3)).max() < 1e-2 and np.array_equal(M[:3], C[:3])
    # one hit: the output link is one heatbath result, in SU(N)
    U = oracle.hot_start(3, dims, 8)
    M1 = oracle.multihit_ext(U, dims, 6.0, 5, 1)
    assert np.abs(M1[3] @ np.conj(np.swapaxes(M1[3], 1, 2)) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(M1[3]) - 1).max() < 1e-12
    # a site's output depends on the seed; different seeds differ
    M2 = oracle.multihit_ext(U, dims, 6.0, 6, 1)
    assert np.abs(M2[3] - M1[3]).max() > 1e-3
    with pytest.raises(ValueError):
        oracle.multihit_ext(U, dims, 6.0, 5, 0)
