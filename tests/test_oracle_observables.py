"""Oracle plaquette / Polyakov loop and ensemble physics, CPU only.

* plaquette (plaquette_cub.cu:96-149,234) and Polyakov loop (polyakovloop.cu:53-99,
  651) against independent lexicographic numpy implementations;
* cold start gives exactly 1 for both;
* both are invariant under a random gauge transformation
  U_mu(x) -> g(x) U_mu(x) g(x+mu)^+ and under a lattice translation;
* heatbath ensembles reproduce the strong-coupling expansion at small beta
  (<P> = beta/4 - beta^3/96 for SU(2), beta/18 + beta^2/216 for SU(3), beta/32 for
  SU(4) at leading order) and the known
  SU(3) Wilson plaquette at beta = 6.0 (about 0.5937 on large volumes).
"""
import numpy as np
import pytest

from test_oracle_math import lex_field, shift

NCS = [2, 3, 4]


def dag(M):
    return np.conj(np.swapaxes(M, -1, -2))


def numpy_plaquette(L, nc):
    V = np.prod(L.shape[1:5])
    sp = tp = 0.0
    for mu in range(3):
        for nu in range(mu + 1, 4):
            P = L[mu] @ shift(L[nu], mu, 1) @ dag(shift(L[mu], nu, 1)) @ dag(L[nu])
            tr = np.trace(P, axis1=-2, axis2=-1).real.sum()
            if nu == 3:
                tp += tr
            else:
                sp += tr
    return np.array([sp, tp]) / (3 * nc * V)


def numpy_polyakov(L, nc):
    T = L.shape[1]
    P = L[3][0]
    for t in range(1, T):
        P = P @ L[3][t]
    return np.trace(P, axis1=-2, axis2=-1).mean() / nc


def eo_field(oracle, L, dims):
    V = int(np.prod(dims))
    nc = L.shape[-1]
    U = np.zeros((4, V, nc, nc), dtype=complex)
    for x3 in range(dims[3]):
        for x2 in range(dims[2]):
            for x1 in range(dims[1]):
                for x0 in range(dims[0]):
                    U[:, oracle.coords_to_eo((x0, x1, x2, x3), dims)] = L[:, x3, x2, x1, x0]
    return U


@pytest.mark.parametrize("nc", NCS)
def test_cold_start_is_one(oracle, nc):
    U = oracle.cold_start(nc, (4, 4, 4, 6))
    assert np.array_equal(oracle.plaquette(U, (4, 4, 4, 6)), [1.0, 1.0])
    assert oracle.polyakov(U, (4, 4, 4, 6)) == 1.0


@pytest.mark.parametrize("nc", NCS)
def test_observables_brute_force(oracle, nc):
    dims = (4, 2, 6, 4)
    U = oracle.hot_start(nc, dims, 31)
    oracle.overrelax_sweep(U, dims)
    L = lex_field(oracle, U, dims)
    assert np.abs(oracle.plaquette(U, dims) - numpy_plaquette(L, nc)).max() < 1e-14
    assert abs(oracle.polyakov(U, dims) - numpy_polyakov(L, nc)) < 1e-14


@pytest.mark.parametrize("nc", NCS)
def test_gauge_and_translation_invariance(oracle, nc):
    dims = (4, 4, 2, 6)
    U = oracle.hot_start(nc, dims, 41)
    L = lex_field(oracle, U, dims)
    g = oracle.hot_start(nc, dims, 42)[0]            # V random SU(N) matrices, eo order
    G = lex_field(oracle, np.stack([g] * 4), dims)[0]
    Lg = np.stack([G @ L[mu] @ dag(shift(G, mu, 1)) for mu in range(4)])
    Ug = eo_field(oracle, Lg, dims)
    p, pg = oracle.plaquette(U, dims), oracle.plaquette(Ug, dims)
    assert np.abs(p - pg).max() < 1e-14
    assert abs(oracle.polyakov(U, dims) - oracle.polyakov(Ug, dims)) < 1e-14
    for mu in range(4):
        Ls = np.stack([shift(L[nu], mu, 1) for nu in range(4)])
        Us = eo_field(oracle, Ls, dims)
        assert np.abs(oracle.plaquette(Us, dims) - p).max() < 1e-14
        assert abs(oracle.polyakov(Us, dims) - oracle.polyakov(U, dims)) < 1e-14


@pytest.mark.parametrize("nc,beta", [(2, 0.4), (3, 0.6), (4, 0.8)])
def test_strong_coupling(oracle, nc, beta):
    dims = (4, 4, 4, 4)
    U = oracle.hot_start(nc, dims, 50 + nc)
    vals = []
    for it in range(160):
        oracle.heatbath_sweep(U, dims, beta, 7, it)
        if it >= 10:
            vals.append(oracle.plaquette(U, dims).mean())
    vals = np.array(vals)
    # <P> = (beta/N^2) <(Re Tr U)^2>_Haar + ...; <(Re Tr U)^2> = 1 for SU(2) (real
    # trace), 1/2 for N >= 3.  SU(2) exact single plaquette: I2(beta)/I1(beta).
    lead = beta / (nc * nc) if nc == 2 else beta / (2 * nc * nc)
    second = {2: -beta ** 3 / 96, 3: beta ** 2 / 216, 4: 0.0}[nc]
    err = 4 * vals.std() / np.sqrt(len(vals) / 2) + 0.1 * lead * (beta / 2) ** 2
    assert abs(vals.mean() - (lead + second)) < err, (vals.mean(), lead + second, err)


@pytest.mark.slow
def test_su3_beta6_plaquette(oracle):
    """SU(3) Wilson action at beta = 6.0: <P> ~ 0.5937 (literature, large volume)."""
    dims = (8, 8, 8, 8)
    U = oracle.cold_start(3, dims)
    vals = []
    for it in range(90):
        oracle.heatbath_sweep(U, dims, 6.0, 3, it)
        for _ in range(4):
            oracle.overrelax_sweep(U, dims)
        oracle.reunitarize(U)
        if it >= 30:
            vals.append(oracle.plaquette(U, dims).mean())
    assert abs(np.mean(vals) - 0.5937) < 0.003
