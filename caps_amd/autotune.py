"""Autotuner of the determinant stage's LU schedule.

For a grid of batch shapes (matrix dimension, number of matrices) it times
the library's batched LU (capsx_lu_time_form: synthetic matrices
1 - 0.3 X/sqrt(N) generated on the GPU, events around the LU only; by default
in the delta form that round-trip blocks take, --pivoted for the pivoted LU)
under the schedules

    levels  1 (64-column steps), 2 (128-column steps, one K = 128 update),
            3 (256-column steps, one K = 256 update) or 4 (512-column steps,
            one K = 512 update)
    tile    trailing-update workgroup tile 128 x 128, 256 x 128 or 64 x 64
    streams 1, 2 or 3 groups of matrices on separate HIP streams
    xcd     trailing-update tile order (1: all tiles of a matrix on one XCD)

first (levels, tile) with one stream, then (streams, xcd) for the best of
those, and writes the fastest schedule per shape to lu_tuning.txt, which the
library reads at start-up (nearest shape in log dimension and log count).

    python -m caps_amd.autotune [--out caps_amd/lu_tuning.txt] [--max-gb 64] [--pivoted]
"""
import argparse
import ctypes
import itertools
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import caps_amd  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "lu_tuning.txt")
DIMS = (240, 600, 1200, 2400, 4800, 9600, 24000)
COUNTS = (1, 4, 16, 128, 640)


def _sig():
    L = caps_amd.lib()
    i = ctypes.c_int
    L.capsx_lu_time.restype = i
    L.capsx_lu_time.argtypes = [i, i, i, i, i, i, i, ctypes.POINTER(ctypes.c_double)]
    L.capsx_lu_time_form.restype = i
    L.capsx_lu_time_form.argtypes = [i, i, i, i, i, i, i, i, ctypes.POINTER(ctypes.c_double)]
    L.capsx_lu_config.restype = i
    L.capsx_lu_config.argtypes = [i, i, i, i]
    L.capsx_lu_config_get.restype = None
    L.capsx_lu_config_get.argtypes = [i, i, ctypes.POINTER(ctypes.c_int)]
    L.capsx_lu_tuning_load.restype = i
    L.capsx_lu_tuning_load.argtypes = [ctypes.c_char_p]
    return L


def lu_time(dim, n, cfg, reps=2, form=1):
    """Milliseconds of one batched LU of n dim x dim matrices under
    cfg = (levels, streams, xcd, tile); form 1: the delta form (1 - M column
    dominant, no pivot search), 0: the pivoted LU."""
    L = _sig()
    ms = ctypes.c_double()
    levels, streams, xcd, tile = cfg
    if L.capsx_lu_time_form(dim, n, reps, levels, streams, xcd, tile, form, ctypes.byref(ms)) != 0:
        raise ValueError("capsx_lu_time_form rejected dim=%d n=%d" % (dim, n))
    return ms.value


def config_for(maxN, n):
    """The schedule the library uses for n blocks of largest dimension maxN."""
    L = _sig()
    c = (ctypes.c_int * 4)()
    L.capsx_lu_config_get(maxN, n, c)
    return tuple(c)


def set_config(levels=0, streams=1, xcd=0, tile=128):
    """Fix the schedule (levels = 0: back to the tuned table)."""
    if _sig().capsx_lu_config(levels, streams, xcd, tile) != 0:
        raise ValueError("invalid LU schedule")


def load_table(path):
    return _sig().capsx_lu_tuning_load(path.encode())


def tune_shape(dim, n, log=print, form=1):
    # the warm-up run of each shape allocates its arena outside the timings
    lu_time(dim, n, (2, 1, 0, 128), reps=1, form=form)
    results = {}
    for levels, tile in itertools.product((1, 2, 3, 4), (128, 256, 64)):
        cfg = (levels, 1, 0, tile)
        results[cfg] = lu_time(dim, n, cfg, form=form)
    lv, _, _, tl = min(results, key=results.get)
    for streams, xcd in itertools.product((1, 2, 3) if n >= 8 else (1,), (0, 1)):
        cfg = (lv, streams, xcd, tl)
        if cfg not in results:
            results[cfg] = lu_time(dim, n, cfg, form=form)
    best = min(results, key=results.get)
    default = results[(2, 1, 0, 128)]
    flops = n * (2.0 / 3 * dim ** 3)
    log("autotune: N=%5d n=%4d best levels=%d streams=%d xcd=%d tile=%d  %.3f ms (%.1f TF/s), "
        "default %.3f ms" % (dim, n, *best, results[best], flops / results[best] / 1e9, default))
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--max-gb", type=float, default=64.0, help="largest batch of matrices to time (GB)")
    ap.add_argument("--pivoted", action="store_true", help="time the pivoted LU instead of the delta form")
    args = ap.parse_args()
    form = 0 if args.pivoted else 1
    set_config(0)
    lines = ["# LU schedule per batch shape, written by caps_amd/autotune.py on %s (%s LU)" %
             (time.strftime("%Y-%m-%d"), "pivoted" if form == 0 else "delta-form"),
             "# maxN n levels streams xcd tile   (# best ms, default ms)"]
    t0 = time.time()
    for dim in DIMS:
        for n in COUNTS:
            if n * dim * dim * 8 > args.max_gb * 1e9:
                continue
            best, res = tune_shape(dim, n, form=form)
            lines.append("%d %d %d %d %d %d   # %.4f %.4f" % (dim, n, *best, res[best], res[(2, 1, 0, 128)]))
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("autotune: wrote %s (%d shapes, %.0f s)" % (args.out, len(lines) - 2, time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
