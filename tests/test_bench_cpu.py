"""bench.py without a GPU: the launch shapes it pins come from the committed rocprof
summaries, the headline summary holds the HBM traffic of exactly the pinned
overrelaxation shape (so `roofline.traffic` is never left empty by a shape the
summary does not cover), and more than one GPU is refused with a message."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_profiled_shapes_parse():
    for name in (bench.PROFILE_32, bench.PROFILE_64):
        sh = bench.profiled_shapes(name)
        assert sh is not None, name
        assert sh["source"] == f"profiles/{name}"
        for key in ("overrelax", "heatbath"):
            c = sh[key]
            assert set(c) == {"block", "brick", "lds", "piped"}
            assert len(c["brick"]) == 3 and c["piped"] in (0, 1, 2) and c["lds"] >= 0


def test_headline_traffic_matches_pinned_shape():
    sh = bench.profiled_shapes(bench.PROFILE_32)
    traffic, src = bench.profile_traffic(sh["overrelax"])
    assert src == sh["source"]
    # between the algorithmic bytes of one launch (9 x 48 B x V/2) and twice them
    alg = 9 * 48 * (32 ** 4 // 2)
    assert alg <= traffic <= 2 * alg


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
