"""Every launch shape the autotuner can pick gives oracle-parity results.

The parity tests elsewhere run the default shape (256-thread linear workgroups).
Here each (N, storage, precision) variant runs one overrelaxation and one heatbath
sweep under every shape that `sun_lattice_set_launch` accepts from a list covering
linear workgroups of 64..512 threads, brick mappings along (x1, x2, x3), LDS
occupancy caps and the software-pipelined kernels, on two geometries whose X0/2 is
4 and 16 sites, and compares each with the FP64 oracle at the usual bars (8-real
FP32: per link, see link_bound).  Shapes a variant cannot run must be refused with
SUN_ERR_INVALID.
"""
import numpy as np
import pytest

import sun_amd
from test_gpu_parity import (BETAS, IDS, MAX_FLIP_OUTLIERS, VARIANTS, flip_report, link_bound, load_config,
                             per_link_diff, thermalised, tol)

pytestmark = pytest.mark.gpu

SHAPES = [  # (block, brick, lds, piped)
    (64, (-1, 0, 0), 0, 0),
    (128, (-1, 0, 0), 16384, 0),
    (512, (-1, 0, 0), 0, 0),
    (0, (1, 1, 1), 0, 0),
    (0, (2, 1, 1), 0, 0),
    (0, (2, 0, 2), 54272, 0),
    (0, (1, 2, 2), 0, 0),
    (0, (2, 2, 2), 0, 0),
    (256, (-1, 0, 0), 0, 1),
    (0, (2, 0, 2), 0, 1),
    (0, (1, 1, 2), 32768, 1),
    (0, (2, 1, 2), 0, 1),
    (256, (-1, 0, 0), 0, 2),
    (0, (2, 0, 2), 16384, 2),
    (0, (1, 1, 2), 0, 2),
    (0, (0, 0, 2), 0, 0),      # x3-deep bricks
    (0, (0, 0, 3), 0, 2),
    (0, (1, 0, 3), 16384, 1),
    (0, (0, 0, 4), 0, 0),
]
DIMS = [(8, 8, 8, 16), (32, 4, 4, 8)]


def accepted_shapes(lat, kernel):
    out = []
    for block, brick, lds, piped in SHAPES:
        try:
            lat.set_launch(kernel, block, brick, lds, piped)
        except sun_amd.SunError as e:
            assert e.code == sun_amd.ERR_INVALID
            continue
        got = lat.get_launch(kernel)
        assert list(got.brick) == list(brick) and got.lds == lds and got.piped == piped
        if brick[0] < 0:
            assert got.block == block
        out.append((block, brick, lds, piped))
    return out


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("dims", DIMS)
def test_every_shape_overrelax_parity(ctx, oracle, nc, storage, prec, dims):
    U0 = thermalised(oracle, nc, dims, storage)
    with sun_amd.Lattice(ctx, nc, dims, BETAS[nc], storage, prec) as lat:
        U = load_config(lat, U0)
        ref = U.copy()
        oracle.overrelax_sweep(ref, dims, storage)
        bound = link_bound(lat, U, ref, storage, prec, lambda A: oracle.overrelax_sweep(A, dims, storage))
        shapes = accepted_shapes(lat, 1)
        assert len(shapes) >= 3
        for block, brick, lds, piped in shapes:
            lat.set_launch(1, block, brick, lds, piped)
            lat.from_host(U0)
            lat.overrelax()
            d = per_link_diff(lat.to_host(), ref)
            assert np.all(d <= bound), (block, brick, lds, piped, float(d.max()))


@pytest.mark.parametrize("nc,storage,prec", VARIANTS, ids=IDS)
@pytest.mark.parametrize("dims", DIMS)
def test_every_shape_heatbath_parity(ctx, oracle, nc, storage, prec, dims):
    beta, seed, sweep = BETAS[nc], 4321, 3
    U0 = thermalised(oracle, nc, dims, storage)
    with sun_amd.Lattice(ctx, nc, dims, beta, storage, prec) as lat:
        U = load_config(lat, U0)
        ref = U.copy()
        margins = oracle.heatbath_sweep_margins(ref, dims, beta, seed, sweep, storage)
        bar = link_bound(lat, U, ref, storage, prec, lambda A: oracle.heatbath_sweep(A, dims, beta, seed, sweep,
                                                                                     storage))
        shapes = accepted_shapes(lat, 0)
        assert len(shapes) >= 3
        for block, brick, lds, piped in shapes:
            lat.set_launch(0, block, brick, lds, piped)
            lat.from_host(U0)
            lat.heatbath(seed, sweep)
            d = per_link_diff(lat.to_host(), ref)
            if prec == 8:
                assert d.max() <= 1e-12, (block, brick, lds, piped)
            else:
                n_out, n_tie, unexplained = flip_report(d, bar, margins, dims)
                print(f"shape {block} {brick} {lds} {piped}: {n_out} outliers, {n_tie} near-tie")
                assert not unexplained and n_out <= MAX_FLIP_OUTLIERS, (block, brick, lds, piped, n_out,
                                                                        unexplained[:10])


def test_piped_refused_without_kernel(ctx):
    with sun_amd.Lattice(ctx, 3, (8, 8, 8, 8), 6.0, 18, 8) as lat:
        with pytest.raises(sun_amd.SunError) as e:
            lat.set_launch(1, 256, (-1, 0, 0), 0, 1)
        assert e.value.code == sun_amd.ERR_INVALID
        with pytest.raises(sun_amd.SunError):
            lat.set_launch(1, 96)  # not whole wavefronts
        with pytest.raises(sun_amd.SunError):
            lat.set_launch(1, 0, (3, 3, 3))  # 4 * 2^9 threads
