"""GPU parity at the full sizes of the benchmark configurations.

* SU(3) 32^4, beta 6.0, FP32 12-real, launched with the shapes of the committed
  32^4 rocprof summary: one heatbath and one overrelaxation sweep from a generated
  field against the oracle's sweeps from the same (decoded) field -- every link
  within 5e-5 except accept/reject flips the oracle's tie margins explain,
  plaquette and Polyakov loop within 2e-6 (monte.cu:196-222, ovr.cu:182-208).
* SU(3) 64^4, beta 6.2, FP32 8-real (2.15 GB of links), the 64^4 summary's shapes:
  one heatbath and one overrelaxation (parity, mu) launch, each checked at 12 000
  random sites of its class against the oracle's single-link update in the staple
  of the decoded field, every other link unchanged bit for bit; one iteration by
  hipGraph replay equal bit for bit to the same iteration from eager launches; the
  link monitor's unitarity deviation within what 8 FP32 reals can hold.
* The benchmark's 64 lattices (volumes 8^4 ... 24^4, beta in [5.7, 6.3], hot and
  cold starts, FP32 12-real): the first batched heatbath and overrelaxation sweep
  against the oracle for one lattice of each volume class (from a Haar start these
  fields are rough, so the bar adds ten times the oracle's own change under a
  one-ulp input perturbation), and two batched iterations bit-identical, lattice by
  lattice, to each lattice run alone.
* Served jobs of 24^4 and 32^4 with 50-60 iterations beside SU(2) and SU(4) jobs:
  each job's observables equal its alone run bit for bit.
"""
import numpy as np
import pytest

import bench
import sun_amd
from test_gpu_parity import MAX_FLIP_OUTLIERS, flip_report, per_link_diff
from test_server import alone, submit

pytestmark = pytest.mark.gpu

TAG_HB = 0x10000000


def test_config2_32x4_sweeps_against_oracle(ctx, oracle):
    dims, beta, seed, sweep = (32, 32, 32, 32), 6.0, 2, 3
    shapes = bench.profiled_shapes(bench.PROFILE_32)
    assert shapes, "the 32^4 profile summary names its launch shapes"
    with sun_amd.Lattice(ctx, 3, dims, beta, 12, 4) as lat:
        bench.pin_shapes(lat, shapes)
        lat.cold_start()
        lat.generate(seed=seed, n_iter=3, n_or=4, reunit_every=1)
        U = lat.to_host()
        lat.heatbath(seed, sweep)
        G = lat.to_host()
        p_hb, l_hb = lat.plaquette(), lat.polyakov()
        lat.overrelax()
        G2 = lat.to_host()
        p_or, l_or = lat.plaquette(), lat.polyakov()
    margins = oracle.heatbath_sweep_margins(U, dims, beta, seed, sweep, 12)  # U becomes the oracle's field
    d = per_link_diff(G, U)
    n_out, n_tie, unexplained = flip_report(d, np.full(d.size, 5e-5), margins, dims)
    print(f"32^4 heatbath: {n_out} of {d.size} links beyond 5e-5 ({n_tie} at near-tie decisions), "
          f"max elsewhere {float(d[d <= 5e-5].max()):.2e}")
    assert not unexplained, unexplained[:10]
    assert n_out <= MAX_FLIP_OUTLIERS, n_out
    assert np.abs(p_hb - oracle.plaquette(U, dims)).max() <= 2e-6
    assert abs(l_hb - oracle.polyakov(U, dims)) <= 2e-6
    ref = G.copy()
    oracle.overrelax_sweep(ref, dims, 12)
    d2 = per_link_diff(G2, ref)
    print(f"32^4 overrelaxation: max link difference {float(d2.max()):.2e}")
    assert d2.max() <= 5e-5
    assert np.abs(p_or - oracle.plaquette(ref, dims)).max() <= 2e-6
    assert abs(l_or - oracle.polyakov(ref, dims)) <= 2e-6


def sampled_bar(ref8, ref, sens):
    """5e-5 plus four times the 8-real round trip of the oracle's link plus ten times
    its change under a one-ulp (2^-23) perturbation of the update's inputs (the 8-real
    decode divides by |e01|^2 + |e02|^2, so some links are ill-conditioned)"""
    return 5e-5 + 4 * np.abs(ref8 - ref).max() + 10 * sens


def check_sampled_launch(oracle, before, after, dims, parity, mu, ids, update):
    """after = before with the (parity, mu) class updated; the sampled sites against
    update(link, F, id) -> oracle's new link; every link outside the class unchanged"""
    V = int(np.prod(dims))
    half = V // 2
    for m in range(4):
        if m != mu:
            assert np.array_equal(after[m], before[m]), m
    other = slice((1 - parity) * half, (2 - parity) * half)
    assert np.array_equal(after[mu, other], before[mu, other])
    rng = np.random.default_rng(1000 + 4 * parity + mu)
    worst, n_bad = 0.0, 0
    for i in ids:
        F = oracle.staple(before, dims, parity, mu, int(i)).conj().T
        u = before[mu, parity * half + i]
        ref = update(u, F, int(i))
        ref8 = oracle.project_storage(ref[None], 8)[0]
        pu = u * (1 + 2.0 ** -23 * rng.standard_normal(u.shape))
        pF = F * (1 + 2.0 ** -23 * rng.standard_normal(F.shape))
        sens = np.abs(oracle.project_storage(update(pu, pF, int(i))[None], 8)[0] - ref8).max()
        d = np.abs(after[mu, parity * half + i] - ref8).max()
        worst = max(worst, float(d))
        n_bad += int(d > sampled_bar(ref8, ref, sens))
    return worst, n_bad


def test_config3_64x4_sampled_sites_against_oracle(ctx, oracle):
    dims, beta, seed = (64, 64, 64, 64), 6.2, 3
    V = int(np.prod(dims))
    shapes = bench.profiled_shapes(bench.PROFILE_64)
    assert shapes, "the 64^4 profile summary names its launch shapes"
    ids = np.sort(np.random.default_rng(64).choice(V // 2, 12000, replace=False))
    with sun_amd.Lattice(ctx, 3, dims, beta, 8, 4) as lat:
        bench.pin_shapes(lat, shapes)
        lat.hot_start(seed)
        lat.generate(seed=seed, n_iter=2, n_or=4, reunit_every=0)
        before = lat.to_host()
        lat.heatbath_sub(seed, 2, 0, 0)
        after = lat.to_host()
        worst, n_bad = check_sampled_launch(
            oracle, before, after, dims, 0, 0, ids,
            lambda u, F, i: oracle.heatbath_link(u, F, beta, seed, c0=i, c2=2, c3=TAG_HB | (0 << 4) | 0))
        print(f"64^4 heatbath (0, 0): max sampled difference {worst:.2e}, {n_bad} beyond the bar")
        assert n_bad == 0
        del before
        lat.overrelax_sub(1, 3)
        after2 = lat.to_host()
        worst, n_bad = check_sampled_launch(oracle, after, after2, dims, 1, 3, ids,
                                            lambda u, F, i: oracle.overrelax_link(u, F))
        print(f"64^4 overrelaxation (1, 3): max sampled difference {worst:.2e}, {n_bad} beyond the bar")
        assert n_bad == 0


def test_config3_64x4_graph_equals_eager_and_unitarity(ctx):
    dims, beta, seed = (64, 64, 64, 64), 6.2, 3
    shapes = bench.profiled_shapes(bench.PROFILE_64)
    res = []
    with sun_amd.Lattice(ctx, 3, dims, beta, 8, 4) as a, sun_amd.Lattice(ctx, 3, dims, beta, 8, 4) as b:
        for lat in (a, b):
            bench.pin_shapes(lat, shapes)
        a.hot_start(seed)
        b.copy_from(a)
        for lat, graph in ((a, True), (b, False)):
            obs = lat.generate(seed=seed, n_iter=1, n_or=4, reunit_every=0, use_graph=graph)
            res.append((obs, lat.link_monitor(), lat.plaquette(), lat.polyakov()))
    (oa, ma, pa, la), (ob, mb, pb, lb) = res
    assert np.array_equal(oa, ob) and np.array_equal(ma, mb) and np.array_equal(pa, pb) and la == lb
    print(f"64^4 after one iteration: plaquette {pa}, unitarity deviation {ma[4]:.2e}, |det - 1| {ma[5]:.2e}, "
          f"mean det {ma[2]:.9f} {ma[3]:.1e}")
    # 8 FP32 reals fix |e20|^2 = |e01|^2 + |e02|^2 - |e10|^2 only to FP32 rounding
    # (~1e-7 absolute), so the rare link whose |e20| is near 0 decodes ~sqrt(1e-7) off
    # SU(3) (reconstruct_12p_8p.h:59-63; 4.1e-3 at worst of 67M links here); the
    # format keeps every link within 1e-2 and the mean determinant at 1
    assert ma[4] <= 1e-2 and ma[5] <= 1e-2
    assert abs(ma[2] - 1) <= 1e-6 and abs(ma[3]) <= 1e-6


def conditioned_bar(U, sweep):
    """5e-5 per link plus ten times the oracle's own change at the link when its input
    U moves by one FP32 ulp (relative 2^-23): a sweep of a rough field (a Haar start,
    or one sweep after it) divides by small SU(2) projections and amplifies the FP32
    rounding of the device's arithmetic by as much"""
    rng = np.random.default_rng(23)
    A = U.copy()
    P = U * (1 + 2.0 ** -23 * rng.standard_normal(U.shape))
    sweep(A)
    sweep(P)
    return 5e-5 + 10 * per_link_diff(P, A)


def bench_batch_specs():
    """the 64 lattices of the benchmark's batched workload (bench.batched_leg)"""
    rng = np.random.default_rng(4)
    vols = [bench.BATCH_VOLUMES[i] for i in rng.integers(0, len(bench.BATCH_VOLUMES), 64)]
    betas = rng.uniform(5.7, 6.3, 64)
    hot = rng.integers(0, 2, 64)
    return [(v, float(b), bool(h), 1000 + i, 4000 + i) for i, (v, b, h) in enumerate(zip(vols, betas, hot))]


def started(ctx, spec):
    dims, beta, hot, start_seed, _ = spec
    lat = sun_amd.Lattice(ctx, 3, dims, beta, 12, 4)
    if hot:
        lat.hot_start(start_seed)
    else:
        lat.cold_start()
    return lat


def test_config4_64_lattices_sweeps_against_oracle(ctx, oracle):
    specs = bench_batch_specs()
    reps = {}
    for i, s in enumerate(specs):
        reps.setdefault(s[0], i)
    lats = [started(ctx, s) for s in specs]
    try:
        U0 = {i: lats[i].to_host() for i in reps.values()}
        b = sun_amd.Batch(ctx, lats, [s[4] for s in specs])
        try:
            b.heatbath([0] * len(specs))
            G = {i: lats[i].to_host() for i in reps.values()}
            b.overrelax()
            G2 = {i: lats[i].to_host() for i in reps.values()}
        finally:
            b.close()
    finally:
        for lat in lats:
            lat.close()
    for dims, i in sorted(reps.items()):
        beta, seed = specs[i][1], specs[i][4]
        hb = lambda A: oracle.heatbath_sweep(A, dims, beta, seed, 0, 12)  # noqa: E731
        orx = lambda A: oracle.overrelax_sweep(A, dims, 12)  # noqa: E731
        bar_hb, bar_or = conditioned_bar(U0[i], hb), conditioned_bar(G[i], orx)
        ref = U0[i]
        margins = oracle.heatbath_sweep_margins(ref, dims, beta, seed, 0, 12)
        d = per_link_diff(G[i], ref)
        n_out, n_tie, unexplained = flip_report(d, bar_hb, margins, dims)
        ref2 = G[i].copy()
        orx(ref2)
        d2 = per_link_diff(G2[i], ref2)
        print(f"batch lattice {i} {dims} beta {beta:.3f} {'hot' if specs[i][2] else 'cold'}: heatbath {n_out} "
              f"outliers ({n_tie} near-tie), max {float(d.max()):.2e}; overrelaxation max {float(d2.max()):.2e} "
              f"(bar up to {float(bar_or.max()):.2e})")
        assert not unexplained and n_out <= MAX_FLIP_OUTLIERS, (dims, n_out, unexplained[:10])
        assert np.all(d2 <= bar_or), dims


def test_config4_64_lattices_batched_equals_alone(ctx):
    specs = bench_batch_specs()
    lats = [started(ctx, s) for s in specs]
    try:
        b = sun_amd.Batch(ctx, lats, [s[4] for s in specs])
        try:
            obs = b.generate(first_sweep=0, n_iter=2, n_or=4, reunit_every=1, use_graph=True)
        finally:
            b.close()
        for i, s in enumerate(specs):
            with started(ctx, s) as one:
                o = one.generate(seed=s[4], n_iter=2, n_or=4, reunit_every=1, first_sweep=0)
                assert np.array_equal(obs[:, i, :], o[:, :2]), i
                assert np.array_equal(lats[i].to_host(), one.to_host()), i
    finally:
        for lat in lats:
            lat.close()


# (nc, dims, beta, storage, precision, hot, n_iter, n_or, reunit_every, seed)
SERVED = [
    (3, (32, 32, 32, 32), 6.0, 12, 4, False, 50, 4, 1, 51),
    (3, (24, 24, 24, 24), 5.8, 12, 4, True, 60, 4, 1, 52),
    (3, (24, 24, 24, 24), 6.1, 8, 4, True, 55, 4, 0, 53),
    (2, (16, 16, 16, 16), 2.4, 18, 4, True, 50, 4, 1, 54),
    (4, (12, 12, 12, 12), 10.8, 18, 4, True, 50, 4, 1, 55),
]


def test_config5_served_large_jobs_equal_alone(ctx):
    with sun_amd.Server(ctx, 8 << 30) as srv:
        srv.start()
        ids = [submit(srv, j) for j in SERVED]
        for jid in ids:
            srv.wait(jid, 120.0)
        srv.stop()
        for jid, job in zip(ids, SERVED):
            assert srv.status(jid).state == sun_amd.JOB_DONE, (job, srv.job_error(jid))
            got = srv.observables(jid, job[6])
            assert np.array_equal(got, alone(ctx, job)), job
