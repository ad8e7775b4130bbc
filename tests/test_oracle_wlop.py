"""Operator-basis Wilson loops of the oracle: the operator fields of
src/wl/calcop_dg_A0.cu (with include/mop/mo0.cuh; 10 slots) and
src/wl/calcop_dg_33.cu (25 slots), and the correlator of wilsonloop_dg_A0.cu.

Pins, none of which restates the oracle's own formulas:
* identity field: every path product is 1, so the slots are the sums of their path
  weights (A0: 1, 2 x8, 0; 33: slot 0 1, slot 1 2 sqrt 2, slot 2 4, slot 3 2 sqrt 2,
  every other combination cancels);
* gauge covariance O(x) -> g(x) O(x) g(x + r mu)^+ (every path runs from x to
  x + r mu);
* the 90-degree rotation about mu: slots 0-8 are invariant, 9-16 pick up i, 17-24
  pick up -1 (the Sigma / Pi / Delta labels of the operator basis); the one exception
  is slot 18, whose weights at calcop_dg_33.cu:524-538 mix the rotation phases (the
  reference's table, reproduced as written);
* reflection of a transverse direction: the Sigma and Delta slots are eigen-operators
  with sign -1 for slots 4, 8, 19, 23 and +1 for the others (slot 18 again excepted);
* abelian flux fields, where every path is a phase fixed by the area it encloses
  (Stokes): MO0(l) for a flux in the (mu, mu+1) plane, and the three all-plus sums
  (slots 1, 2, 3) for a flux carried by the links along mu;
* the correlator: slot pair (0, 0) summed over mu and divided by 3 N V is W(r, T) of
  or_wilson_loops (pinned in test_oracle_meas.py); at T = 0 it is
  sum_x Re Tr(O_i O_j^+) (numpy); with unit temporal links it is
  sum_x Re Tr(O_i(x) O_j(x + T e3)^+); it is gauge invariant.
"""
import numpy as np
import pytest

from test_oracle_meas import gauge_transform, haar

DIMS = (4, 4, 4, 6)
V = int(np.prod(DIMS))
RT8 = 2.0 * np.sqrt(2.0)


@pytest.fixture(scope="module")
def field(oracle):
    U = oracle.hot_start(3, DIMS, 5)
    for it in range(2):
        oracle.heatbath_sweep(U, DIMS, 5.0, 9, it)
    return U


def coords_of(oracle, dims):
    v = int(np.prod(dims))
    return np.array([oracle.eo_to_coords(s % (v // 2), s // (v // 2), dims) for s in range(v)])


def eo_of(x, dims):
    v = int(np.prod(dims))
    lex = ((x[3] * dims[2] + x[2]) * dims[1] + x[1]) * dims[0] + x[0]
    return lex // 2 + (sum(x) & 1) * (v // 2)


def rotate(oracle, U, dims, mu):
    """the field rotated by 90 degrees about mu: e_{mu+1} -> e_{mu+2} -> -e_{mu+1};
    returns it and the site map s -> eo(R x)"""
    d2, d3 = (mu + 1) % 3, (mu + 2) % 3
    Ur, smap = np.empty_like(U), np.empty(V, int)
    for s, x in enumerate(coords_of(oracle, dims)):
        y = list(x)
        y[d3], y[d2] = x[d2], (-x[d3]) % dims[d2]
        sy = eo_of(y, dims)
        smap[s] = sy
        Ur[mu, sy], Ur[3, sy], Ur[d3, sy] = U[mu, s], U[3, s], U[d2, s]
        y[d2] = (y[d2] - 1) % dims[d2]
        Ur[d2, eo_of(y, dims)] = U[d3, s].conj().T
    return Ur, smap


def reflect(oracle, U, dims, d):
    """the field reflected in direction d (x_d -> -x_d)"""
    Ur, smap = np.empty_like(U), np.empty(V, int)
    for s, x in enumerate(coords_of(oracle, dims)):
        y = list(x)
        y[d] = (-x[d]) % dims[d]
        sy = eo_of(y, dims)
        smap[s] = sy
        for nu in range(4):
            if nu != d:
                Ur[nu, sy] = U[nu, s]
        y[d] = (y[d] - 1) % dims[d]
        Ur[d, eo_of(y, dims)] = U[d, s].conj().T
    return Ur, smap


def eigen(A, B):
    """lambda with A = lambda B, and the relative residual"""
    lam = np.vdot(B, A) / np.vdot(B, B).real
    return lam, np.abs(A - lam * B).max() / np.abs(B).max()


@pytest.mark.parametrize("family", [0, 1], ids=["A0", "33"])
def test_identity_field(oracle, family):
    I3 = np.eye(3)
    for r in (1, 2, 3):
        O = oracle.wlop_fields(oracle.cold_start(3, DIMS), DIMS, family, r, r % 3)
        if family == 0:
            want = [I3] + [2 * I3] * 8 + [0 * I3]
        else:
            want = [I3, RT8 * I3, 4 * I3, RT8 * I3] + [0 * I3] * 21
        for k, w in enumerate(want):
            assert np.abs(O[k] - w).max() < 1e-15, (r, k)


@pytest.mark.parametrize("family", [0, 1], ids=["A0", "33"])
def test_gauge_covariance(oracle, field, family):
    rng = np.random.default_rng(21)
    G = haar(rng, 3, (V,))
    Ug = gauge_transform(field, DIMS, G)
    xs = coords_of(oracle, DIMS)
    for r, mu in ((2, 0), (3, 1), (4, 2)):
        O = oracle.wlop_fields(field, DIMS, family, r, mu)
        Og = oracle.wlop_fields(Ug, DIMS, family, r, mu)
        far = []
        for x in xs:
            y = list(x)
            y[mu] = (y[mu] + r) % DIMS[mu]
            far.append(eo_of(y, DIMS))
        want = G[None] @ O @ G[np.array(far)][None].conj().swapaxes(-1, -2)
        assert np.abs(Og - want).max() < 1e-13, (r, mu)
        assert np.abs(O[1:9]).max() > 0.1


@pytest.mark.parametrize("family", [0, 1], ids=["A0", "33"])
def test_rotation_about_the_axis(oracle, field, family):
    for r in (2, 3):
        for mu in range(3):
            Ur, smap = rotate(oracle, field, DIMS, mu)
            O = oracle.wlop_fields(field, DIMS, family, r, mu)
            Or = oracle.wlop_fields(Ur, DIMS, family, r, mu)
            for k in range(O.shape[0]):
                if family == 0 and k == 9:
                    assert not Or[k].any() and not O[k].any()
                    continue
                lam, res = eigen(Or[k][smap], O[k])
                if family == 1 and k == 18:
                    assert res > 0.1, (r, mu)
                    continue
                want = 1 if (family == 0 or k <= 8) else (1j if k <= 16 else -1)
                assert abs(lam - want) < 1e-13 and res < 1e-13, (r, mu, k, lam, res)


def test_reflection_of_a_transverse_direction(oracle, field):
    odd = {4, 8, 19, 23}
    for r in (2, 3):
        for mu in range(3):
            for d in ((mu + 1) % 3, (mu + 2) % 3):
                Ur, smap = reflect(oracle, field, DIMS, d)
                O = oracle.wlop_fields(field, DIMS, 1, r, mu)
                Or = oracle.wlop_fields(Ur, DIMS, 1, r, mu)
                for k in list(range(9)) + [17] + list(range(19, 25)):
                    lam, res = eigen(Or[k][smap], O[k])
                    assert abs(lam - (-1 if k in odd else 1)) < 1e-13 and res < 1e-13, (r, mu, d, k)
                A0 = oracle.wlop_fields(field, DIMS, 0, r, mu)
                A0r = oracle.wlop_fields(Ur, DIMS, 0, r, mu)
                assert np.abs(A0r[:, smap] - A0).max() < 1e-13


def flux_field(oracle, dims, carrier, coord, n):
    """U(1) flux: the links along `carrier` are diag(e^{i a y}, e^{-i a y}, 1) with
    y = x_coord and a = 2 pi n / X_coord, every other link 1"""
    U = oracle.cold_start(3, dims)
    a = 2 * np.pi * n / dims[coord]
    for s, x in enumerate(coords_of(oracle, dims)):
        U[carrier, s] = np.diag([np.exp(1j * a * x[coord]), np.exp(-1j * a * x[coord]), 1.0])
    return U, a


def test_mo0_heights_by_stokes(oracle):
    dims = (6, 6, 4, 4)
    for mu, r in ((0, 2), (1, 3), (2, 1)):
        d2 = (mu + 1) % 3
        U, a = flux_field(oracle, dims, d2, mu, 1)
        O = oracle.wlop_fields(U, dims, 0, r, mu)
        for l in range(1, 9):
            # the two staples in the flux plane enclose l r plaquettes each, with
            # opposite orientation; the other two are 1
            c = 1 + np.cos(a * l * r)
            assert np.abs(O[l] - np.diag([c, c, 2.0])).max() < 1e-14, (mu, r, l)


def test_33_plus_sums_by_stokes(oracle):
    dims = (6, 6, 6, 4)
    xs = coords_of(oracle, dims)
    for mu, r in ((0, 2), (1, 3), (2, 4), (0, 5)):
        d2 = (mu + 1) % 3
        U, a = flux_field(oracle, dims, mu, d2, 1)
        O = oracle.wlop_fields(U, dims, 1, r, mu)
        h = r // 2
        ph = np.exp(1j * a * xs[:, d2] * r)
        # every path is e^{i a y0 r} times e^{+-i a h} (a staple displaced one step
        # across the flux: families I, III) or 1 (displaced along the flux-free
        # direction); family II always crosses once, half of its paths each way
        s13 = (4 * np.cos(a * h) + 4) / RT8
        s2 = 16 * np.cos(a * h) / 4
        for slot, s, s22 in ((1, s13, 8 / RT8), (2, s2, 4.0), (3, s13, 8 / RT8)):
            assert np.abs(O[slot][:, 0, 0] - s * ph).max() < 1e-13, (mu, r, slot)
            assert np.abs(O[slot][:, 1, 1] - s * ph.conj()).max() < 1e-13, (mu, r, slot)
            assert np.abs(O[slot][:, 2, 2] - s22).max() < 1e-13, (mu, r, slot)


@pytest.mark.parametrize("family", [0, 1], ids=["A0", "33"])
def test_line_pair_is_the_wilson_loop(oracle, field, family):
    rmax, tmax = 3, 4
    W = oracle.wilson_loops(field, DIMS, rmax, tmax)
    got = oracle.wilson_loops_ops(field, DIMS, rmax, tmax, family)
    assert got.shape == (rmax - 1, oracle.WLOP_NOPS[family] ** 2, tmax + 1)
    assert np.abs(got[:, 0, :] - W[2:].real).max() < 1e-14


def test_correlator_at_zero_extent_and_unit_time_links(oracle, field):
    U = field.copy()
    O = oracle.wlop_fields(U, DIMS, 0, 2, 1)
    nop = O.shape[0]
    got = oracle.wlop_correlate(U, DIMS, O, 2, 1, 3)
    want0 = np.einsum("iaxy,jaxy->ij", O, O.conj()).real.reshape(-1)
    assert np.abs(got[:, 0] - want0).max() < 1e-11
    assert np.abs(got[:, 0].reshape(nop, nop) - got[:, 0].reshape(nop, nop).T).max() < 1e-11
    # unit temporal links: t0 = t1 = 1, so W_ij(T) = sum_x Re Tr(O_i(x) O_j(x + T e3)^+)
    U[3] = np.eye(3)
    O = oracle.wlop_fields(U, DIMS, 1, 3, 2)
    got = oracle.wlop_correlate(U, DIMS, O, 3, 2, 4)
    xs = coords_of(oracle, DIMS)
    for T in range(5):
        up = np.array([eo_of([x[0], x[1], x[2], (x[3] + T) % DIMS[3]], DIMS) for x in xs])
        want = np.einsum("iaxy,jaxy->ij", O, O[:, up].conj()).real.reshape(-1)
        assert np.abs(got[:, T] - want).max() < 1e-11, T


def test_correlator_gauge_invariant(oracle, field):
    rng = np.random.default_rng(22)
    Ug = gauge_transform(field, DIMS, haar(rng, 3, (V,)))
    for family in (0, 1):
        a = oracle.wlop_correlate(field, DIMS, oracle.wlop_fields(field, DIMS, family, 2, 0), 2, 0, 3)
        b = oracle.wlop_correlate(Ug, DIMS, oracle.wlop_fields(Ug, DIMS, family, 2, 0), 2, 0, 3)
        assert np.abs(a - b).max() < 1e-10 * np.abs(a).max()


def test_bad_arguments(oracle, field):
    for bad in ((0, 1, 3), (0, 0, 0), (1, 5, 0), (0, 2, -1)):
        fam, r, mu = bad
        with pytest.raises(ValueError):
            oracle.wlop_fields(field, DIMS, fam, r, mu)
