"""GPU tests of the generation driver, hipGraph replay, batching and tuning.

* sun_generate on the reference's own small case (SU(3) 8^4, beta=6.0, Haar hot
  start seed 1, FP64 18-real, 10 x (1 heatbath + 4 overrelaxation), plaquette every
  iteration) runs free beside the FP64 oracle's chain and agrees link by link;
* a captured-and-replayed hipGraph produces bit-identical links and observables to
  eager launches, on every replay;
* a lattice gives bit-identical results alone and inside a batch of lattices of
  other volumes and couplings (the serving determinism requirement);
* the autotuner's block size never changes results.
"""
import numpy as np
import pytest

import sun_amd

pytestmark = pytest.mark.gpu


def test_generate_config1_chain(ctx, oracle):
    """The reference's own small case, run free on both sides: SU(3) 8^4, beta 6.0,
    Haar start from seed 1, FP64 18-real, 10 iterations of 1 heatbath + 4
    overrelaxation sweeps through one sun_generate call (hipGraph replay), against
    the oracle's chain from the same start.  The device runs the oracle's FP64
    operations in the oracle's order with the same portable transcendentals, so the
    links agree to 1e-12 after 10 chaotic iterations (a sweep amplifies a 1e-15
    input difference ~100x) and every iteration's plaquette and Polyakov loop to
    1e-13."""
    dims, beta, seed, n_iter = (8, 8, 8, 8), 6.0, 1, 10
    U = oracle.hot_start(3, dims, seed)
    with sun_amd.Lattice(ctx, 3, dims, beta, 18, 8) as lat:
        lat.hot_start(seed)
        assert np.abs(lat.to_host() - U).max() <= 1e-15
        obs = lat.generate(seed=seed, n_iter=n_iter, n_or=4, reunit_every=0, first_sweep=0)
        G = lat.to_host()
    for it in range(n_iter):
        oracle.heatbath_sweep(U, dims, beta, seed, it)
        for _ in range(4):
            oracle.overrelax_sweep(U, dims)
        assert np.abs(obs[it, :2] - oracle.plaquette(U, dims)).max() <= 1e-13, it
        assert abs(complex(obs[it, 2], obs[it, 3]) - oracle.polyakov(U, dims)) <= 1e-13, it
    d = float(np.abs(G - U).max())
    print(f"config 1 chain: max link difference after {n_iter} iterations {d:.3e}")
    assert d <= 1e-12


def mean_err(x, nblocks=5):
    """mean and batch-means standard error (the chain is autocorrelated)"""
    x = np.asarray(x)
    n = len(x) // nblocks * nblocks
    b = x[len(x) - n:].reshape(nblocks, -1).mean(axis=1)
    return float(b.mean()), float(b.std(ddof=1) / np.sqrt(nblocks))


@pytest.mark.parametrize("nc,beta,storage", [(3, 6.0, 12), (2, 2.3, 18), (4, 11.0, 18)])
def test_ensemble_agrees_with_oracle(ctx, oracle, nc, beta, storage):
    """FP32 GPU chain vs FP64 oracle chain (same seeds, diverging after a few
    iterations): the means over iterations 10..49 of the plaquette and of the
    Polyakov loop's modulus agree within 4 sigma of their batch-means errors
    (plaquette_cub.cu:234, polyakovloop.cu:651).  6^4 keeps |P| large enough (the
    loop is 6 links long) to be measured against its error."""
    dims = (6, 6, 6, 6)
    n_iter, burn = 50, 10
    with sun_amd.Lattice(ctx, nc, dims, beta, storage, 4) as lat:
        lat.cold_start()
        obs = lat.generate(seed=3, n_iter=n_iter, n_or=4, reunit_every=1)
    g = obs[burn:, :2].mean(axis=1)
    gl = np.hypot(obs[burn:, 2], obs[burn:, 3])
    U = oracle.cold_start(nc, dims)
    o, ol = [], []
    for it in range(n_iter):
        oracle.heatbath_sweep(U, dims, beta, 3, it)
        for _ in range(4):
            oracle.overrelax_sweep(U, dims)
        oracle.reunitarize(U)
        o.append(oracle.plaquette(U, dims).mean())
        ol.append(abs(oracle.polyakov(U, dims)))
    (mg, eg), (mo, eo) = mean_err(g), mean_err(o[burn:])
    (mgl, egl), (mol, eol) = mean_err(gl), mean_err(ol[burn:])
    print(f"su{nc} s{storage}: plaquette {mg:.6f}({eg:.1e}) vs oracle {mo:.6f}({eo:.1e}); "
          f"|Polyakov| {mgl:.5f}({egl:.1e}) vs {mol:.5f}({eol:.1e})")
    assert abs(mg - mo) <= 4 * np.hypot(eg, eo) + 1e-6, (mg, mo, eg, eo)
    assert abs(mgl - mol) <= 4 * np.hypot(egl, eol) + 1e-6, (mgl, mol, egl, eol)


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (16, 16, 16, 16)])
def test_precision_lockstep(ctx, dims):
    """Precision target, measured the only way a chaotic chain allows: each of 20
    iterations (HB + 4 OR + reunit, beta 6.0, from a cold start thermalised 5
    iterations) run in FP32 12-real from the FP64 18-real chain's field of the same
    iteration (converted on the device): plaquette within 1e-6 of the FP64
    iteration's.  The Polyakov loop averages only X*Y*Z lines of T links each (512
    lines at 8^4), so its FP32 error is larger: stated bar 1e-5 (measured 1.6e-6 at
    8^4)."""
    seed = 7
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as l64, sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as l32:
        l64.cold_start()
        l64.generate(seed=seed, n_iter=5, n_or=4, reunit_every=1)
        dp, dl = [], []
        for it in range(5, 25):
            l32.convert_from(l64)
            o32 = l32.generate(seed=seed, n_iter=1, n_or=4, reunit_every=1, first_sweep=it, use_graph=False)
            o64 = l64.generate(seed=seed, n_iter=1, n_or=4, reunit_every=1, first_sweep=it, use_graph=False)
            dp.append(float(np.abs(o32[0, :2] - o64[0, :2]).max()))
            dl.append(abs(complex(o32[0, 2], o32[0, 3]) - complex(o64[0, 2], o64[0, 3])))
    print(f"lockstep {dims}: plaquette max {max(dp):.2e} mean {np.mean(dp):.2e}; Polyakov max {max(dl):.2e}")
    assert max(dp) <= 1e-6 and max(dl) <= 1e-5, (dp, dl)


@pytest.mark.parametrize("storage,prec", [(12, 4), (18, 8), (8, 4)])
def test_graph_replay_bitwise_equals_eager(ctx, storage, prec):
    dims = (8, 8, 8, 12)
    out = []
    for use_graph in (False, True):
        with sun_amd.Lattice(ctx, 3, dims, 5.9, storage, prec) as lat:
            lat.hot_start(3)
            obs = lat.generate(seed=9, n_iter=4, n_or=2, reunit_every=2, use_graph=use_graph)
            out.append((obs, lat.to_host()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_generate_continues_chain(ctx):
    """two calls of n iterations == one call of 2n (sweep counter and RNG continue)"""
    dims = (4, 4, 8, 8)
    res = []
    for split in (False, True):
        with sun_amd.Lattice(ctx, 3, dims, 6.0, 12, 4) as lat:
            lat.cold_start()
            if split:
                a = lat.generate(seed=5, n_iter=3, first_sweep=0)
                b = lat.generate(seed=5, n_iter=3, first_sweep=3)
                obs = np.concatenate([a, b])
            else:
                obs = lat.generate(seed=5, n_iter=6, first_sweep=0)
            res.append((obs, lat.to_host()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("storage,prec", [(12, 4), (18, 8)])
def test_batch_bitwise_equals_alone(ctx, storage, prec):
    specs = [((8, 8, 8, 8), 5.8, 11, True), ((4, 6, 8, 10), 6.1, 12, False), ((6, 6, 6, 12), 6.3, 13, True)]
    alone = []
    for dims, beta, seed, hot in specs:
        with sun_amd.Lattice(ctx, 3, dims, beta, storage, prec) as lat:
            lat.hot_start(seed) if hot else lat.cold_start()
            obs = lat.generate(seed=seed, n_iter=3, n_or=4, reunit_every=1, first_sweep=0, use_graph=False)
            alone.append((obs[:, :2], lat.to_host()))
    lats = []
    for dims, beta, seed, hot in specs:
        lat = sun_amd.Lattice(ctx, 3, dims, beta, storage, prec)
        lat.hot_start(seed) if hot else lat.cold_start()
        lats.append(lat)
    b = sun_amd.Batch(ctx, lats, [s[2] for s in specs])
    obs = b.generate(first_sweep=0, n_iter=3, n_or=4, reunit_every=1, use_graph=True)
    for i, lat in enumerate(lats):
        assert np.array_equal(obs[:, i, :], alone[i][0])
        assert np.array_equal(lat.to_host(), alone[i][1])
    b.close()
    for lat in lats:
        lat.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_chunks_bitwise_equal_alone(ctx, use_graph):
    """A batch runs in chunks of consecutive lattices with at most a third of 256 MB of
    links, three chunks at a time on three streams (chunk i on stream i % 3); a lattice
    larger than that is a chunk of its own.  Here 8^4 + 16^4, a 302 MB 32^3x48 lattice,
    12^4 + 8^4 + 24^4, then two more 24^4 (64 MB each): five chunks, so two of the
    streams take two chunks each.  Every lattice's observables and links equal its alone
    run bit for bit, with the iterations captured in a graph and launched eagerly."""
    specs = [((8, 8, 8, 8), 5.8, 41), ((16, 16, 16, 16), 6.0, 42), ((32, 32, 32, 48), 6.1, 43),
             ((12, 12, 12, 12), 6.2, 44), ((8, 8, 8, 8), 6.3, 45), ((24, 24, 24, 24), 5.9, 46),
             ((24, 24, 24, 24), 6.05, 47), ((24, 24, 24, 24), 6.15, 48)]
    alone = []
    for dims, beta, seed in specs:
        with sun_amd.Lattice(ctx, 3, dims, beta, 12, 4) as lat:
            lat.hot_start(seed)
            obs = lat.generate(seed=seed, n_iter=2, n_or=4, reunit_every=1, first_sweep=0, use_graph=False)
            alone.append((obs[:, :2], lat.to_host()))
    lats = []
    for dims, beta, seed in specs:
        lat = sun_amd.Lattice(ctx, 3, dims, beta, 12, 4)
        lat.hot_start(seed)
        lats.append(lat)
    b = sun_amd.Batch(ctx, lats, [s[2] for s in specs])
    obs = b.generate(first_sweep=0, n_iter=2, n_or=4, reunit_every=1, use_graph=use_graph)
    for i, lat in enumerate(lats):
        assert np.array_equal(obs[:, i, :], alone[i][0]), i
        assert np.array_equal(lat.to_host(), alone[i][1]), i
    b.close()
    for lat in lats:
        lat.close()


@pytest.mark.parametrize("storage", [12, 8])
def test_batch_launch_shapes_bitwise_equal(ctx, storage):
    """Every batched launch shape -- plain, direction- and half-direction-pipelined
    staple, linear or per-lattice brick order, LDS occupancy caps -- gives the same
    bits, on lattices whose X0/2 is 4 (bricks of 4x4x4 rows), 5 (no brick: linear)
    and 16 (4x1x4 rows)."""
    specs = [((8, 8, 8, 8), 5.8, 31), ((10, 6, 4, 8), 6.1, 32), ((32, 4, 4, 8), 6.3, 33)]
    shapes = [(0, 0, 0, 0), (0, 0, 1, 0), (1, 2, 0, 0), (2, 1, 1, 32768), (2, 2, 1, 54272), (0, 1, 1, 0),
              (1, 1, 0, 16384)]
    res = {}
    for hb_mode, or_mode, brick, lds in shapes:
        lats = []
        for dims, beta, seed in specs:
            lat = sun_amd.Lattice(ctx, 3, dims, beta, storage, 4)
            lat.hot_start(seed)
            lats.append(lat)
        b = sun_amd.Batch(ctx, lats, [s[2] for s in specs])
        b.set_launch(0, hb_mode, brick, lds)
        b.set_launch(1, or_mode, brick, lds)
        assert (b.get_launch(0), b.get_launch(1)) == ((hb_mode, brick, lds), (or_mode, brick, lds))
        obs = b.generate(first_sweep=0, n_iter=2, n_or=3, reunit_every=1, use_graph=True)
        res[(hb_mode, or_mode, brick, lds)] = (obs, [l.to_host() for l in lats])
        b.close()
        for lat in lats:
            lat.close()
    ref_obs, ref_links = res.pop((0, 0, 0, 0))
    for key, (obs, links) in res.items():
        assert np.array_equal(obs, ref_obs), key
        for a, r in zip(links, ref_links):
            assert np.array_equal(a, r), key


def test_batch_piped_refused_without_kernel(ctx):
    lats = [sun_amd.Lattice(ctx, 3, (4, 4, 4, 4), 6.0, 18, 4), sun_amd.Lattice(ctx, 2, (4, 4, 4, 4), 2.3, 18, 4)]
    for lat in lats:
        b = sun_amd.Batch(ctx, [lat], [1])
        assert b.get_launch(0)[0] == 0 and b.get_launch(1)[0] == 0
        b.set_launch(1, 0, 1, 16384)  # bricks and LDS caps need no pipelined kernel
        assert b.get_launch(1) == (0, 1, 16384)
        for kernel in (0, 1):
            with pytest.raises(sun_amd.SunError) as e:
                b.set_launch(kernel, 1)
            assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError):
            b.set_launch(2, 0)
        with pytest.raises(sun_amd.SunError):
            b.set_launch(0, 0, 2)
        with pytest.raises(sun_amd.SunError):
            b.set_launch(0, 0, 0, 65537)
        b.close()
        lat.close()


def test_batch_step_api_matches_generate(ctx):
    specs = [((8, 8, 8, 8), 6.0, 21), ((4, 4, 4, 8), 5.7, 22)]
    lats = [sun_amd.Lattice(ctx, 3, d, b, 12, 4) for d, b, _ in specs]
    for l, (_, _, s) in zip(lats, specs):
        l.hot_start(s)
    b = sun_amd.Batch(ctx, lats, [s for _, _, s in specs])
    b.heatbath([0, 0])
    for _ in range(2):
        b.overrelax()
    b.reunitarize()
    p = b.plaquette()
    ref = []
    for (d, beta, s) in specs:
        with sun_amd.Lattice(ctx, 3, d, beta, 12, 4) as l:
            l.hot_start(s)
            l.heatbath(s, 0)
            l.overrelax()
            l.overrelax()
            l.reunitarize()
            ref.append(l.plaquette())
    assert np.array_equal(p, np.array(ref))
    b.close()
    for l in lats:
        l.close()


def test_autotuned_block_size_does_not_change_results(oracle):
    res = []
    for tune in (0, 1):
        c = sun_amd.Context(0, 0, tune)
        with sun_amd.Lattice(c, 3, (8, 8, 8, 8), 6.0, 12, 4) as lat:
            lat.hot_start(2)
            lat.heatbath(2, 0)
            lat.overrelax()
            res.append(lat.to_host())
        c.close()
    assert np.array_equal(res[0], res[1])


def test_errors_are_reported_not_fatal(ctx):
    with pytest.raises(sun_amd.SunError, match="even"):
        sun_amd.Lattice(ctx, 3, (5, 4, 4, 4), 6.0, 18, 8)
    with pytest.raises(sun_amd.SunError, match="SU\\(3\\) only"):
        sun_amd.Lattice(ctx, 4, (4, 4, 4, 4), 6.0, 12, 4)
    with pytest.raises(sun_amd.SunError, match="nc=5"):
        sun_amd.Lattice(ctx, 5, (4, 4, 4, 4), 6.0, 18, 4)
    # the context keeps working after errors
    with sun_amd.Lattice(ctx, 3, (4, 4, 4, 4), 6.0, 18, 8) as lat:
        lat.cold_start()
        assert lat.plaquette()[0] == 1.0


def test_cpp_facade_program_matches_c_abi(ctx, tmp_path):
    """A CULQCD-style C++ program (tests/host/facade_demo.cpp against include/sun.hpp)
    produces the same bits as the same run through the C ABI from Python: the
    generation, the link monitors and the measurement chain on normal-order arrays
    (whose host copy is lexicographic)."""
    import os
    import subprocess
    from sun_amd.build import FACADE_DEMO
    assert os.path.exists(FACADE_DEMO), "facade demo not built: run __graft_entry__.build()"
    out, lex = tmp_path / "links.bin", tmp_path / "lex.bin"
    dims, seed, iters = (8, 8, 8, 8), 1, 3
    r = subprocess.run([FACADE_DEMO, *map(str, dims), str(seed), str(iters), str(out), str(lex)],
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    mon = [ln for ln in lines if ln.startswith("monitor ")]
    wl = [ln for ln in lines if ln.startswith("wloop ")]
    rows = np.array([[float(v) for v in ln.split()[1:]] for ln in lines if ln[0].isdigit()])
    links = np.fromfile(out, dtype=np.float64).view(np.complex128).reshape(4, -1, 3, 3)
    lexf = np.fromfile(lex, dtype=np.float64).view(np.complex128).reshape(4, -1, 3, 3)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as lat, sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as mh:
        lat.hot_start(seed)
        obs = lat.generate(seed=seed, n_iter=iters, n_or=4, reunit_every=0, use_graph=False)
        G = lat.to_host()
        m = lat.link_monitor()
        lat.ape_smear(0.2, 3, 10, 1e-10)
        lat.multihit(mh, seed + 1, 5)
        W = mh.wilson_loops(2, 3)
        H = mh.to_host()
    assert np.array_equal(rows, obs)
    assert np.array_equal(links, G)
    assert len(mon) == 1 and [float(v) for v in mon[0].split()[1:]] == list(m[:4])
    assert len(wl) == 12
    for ln in wl:
        _, rr, tt, re, im = ln.split()
        assert complex(float(re), float(im)) == W[int(rr), int(tt)]
    # lexicographic host order of the normal-order array
    V = 8 ** 4
    for lexi in (0, 1, 9, 77, 4095):
        x = [lexi % 8, (lexi // 8) % 8, (lexi // 64) % 8, lexi // 512]
        eo = lexi // 2 + (sum(x) & 1) * (V // 2)
        assert np.array_equal(lexf[:, lexi], H[:, eo])


@pytest.mark.parametrize("src,dst", [((18, 8), (12, 4)), ((12, 4), (18, 8)), ((18, 8), (8, 4)), ((8, 4), (18, 4))])
def test_lattice_convert_equals_host_round_trip(ctx, src, dst):
    """sun_lattice_convert (device to device, any storage and precision) gives the
    bits of the host round trip to_host -> from_host, and refuses other extents or N"""
    dims = (4, 6, 4, 8)
    with sun_amd.Lattice(ctx, 3, dims, 6.0, *src) as a, sun_amd.Lattice(ctx, 3, dims, 6.0, *dst) as b, \
            sun_amd.Lattice(ctx, 3, dims, 6.0, *dst) as c:
        a.hot_start(12)
        a.generate(seed=12, n_iter=2, n_or=2, reunit_every=1)
        b.convert_from(a)
        c.from_host(a.to_host())
        assert np.array_equal(b.to_host(), c.to_host())
    with sun_amd.Lattice(ctx, 3, dims, 6.0, 18, 8) as a, sun_amd.Lattice(ctx, 3, (4, 6, 4, 6), 6.0, 18, 8) as b, \
            sun_amd.Lattice(ctx, 2, dims, 2.3, 18, 8) as c:
        for other in (b, c):
            with pytest.raises(sun_amd.SunError) as e:
                other.convert_from(a)
            assert e.value.code == sun_amd.ERR_INVALID
