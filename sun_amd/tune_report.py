"""Run the launch-shape autotuner on one lattice and print every candidate it times.

    python -m sun_amd.tune_report 32 32 32 32 --nc 3 --storage 12 --precision 4

The autotuner (heatbath and overrelaxation, plain and software-pipelined kernels,
linear and brick workgroup mappings, LDS occupancy caps) writes one line per
candidate to stderr and its choice to the tune cache; the chosen shapes are printed
last.  Under `rocprofv3 --pmc ...` every candidate is 9 consecutive dispatches (one
warm-up, then mu = 0..3 twice), in the order of the printed lines.
"""
import argparse
import os
import sys
import tempfile


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("dims", type=int, nargs=4)
    ap.add_argument("--nc", type=int, default=3)
    ap.add_argument("--storage", type=int, default=12)
    ap.add_argument("--precision", type=int, default=4)
    ap.add_argument("--beta", type=float, default=6.0)
    args = ap.parse_args(argv)
    # a fresh cache directory, so the tuner runs instead of reading an earlier choice
    os.environ["SUN_RESOURCE_PATH"] = tempfile.mkdtemp(prefix="sun_tune_")
    import sun_amd
    ctx = sun_amd.Context(0, 2, 1)
    with sun_amd.Lattice(ctx, args.nc, args.dims, args.beta, args.storage, args.precision) as lat:
        if args.storage == 8:
            lat.hot_start(1)
        else:
            lat.cold_start()
        lat.heatbath(1, 0)
        print("heatbath:", lat.get_launch(0).as_dict())
        print("overrelaxation:", lat.get_launch(1).as_dict())
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
