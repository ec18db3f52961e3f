"""Autotuner of the half-length DCT products' tile shapes (run on an MI355X).

For every folded grid shape in SHAPES, both precisions (FP32 and split-BF16) and the four product
modes, every valid tile shape of gp_kernels.hpp (kDctTiles) is timed with mp_gp_dct_tile_time at
one and at two evaluations per launch chain (the evaluation chain batches the current and the
look-ahead evaluations), and the tile with the smallest summed time wins.  The winners are written
as csrc/dct_tiles.inc, which the library compiles in, and a summary table is printed (markdown).

    python -m myplacement_amd.autotune [--reps 40] [--out myplacement_amd/csrc/dct_tiles.inc]

The table is data, not code: a shape it does not name takes tile 0 (the eight-way contraction split
of 32 x 32 tiles).  Results depend only on a tile's contraction split s, so choosing among tiles
changes speed and, between different s, the rounding of the FP32 sums, never the algorithm.
"""

from __future__ import annotations

import argparse
import ctypes as C
import datetime
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from myplacement_amd import backend as hb  # noqa: E402

MODES = ("fold-cols", "fold-rows", "mirror-cols", "mirror-rows")
_SHAPES = ((32, 32, 1, 8), (32, 32, 1, 4), (16, 32, 1, 4), (32, 32, 1, 2), (32, 32, 1, 1), (32, 32, 2, 1), (16, 32, 2, 1),
           (16, 16, 4, 1))
TILES = tuple(t + (f,) for f in (0, 1) for t in _SHAPES)  # (tr, tc, w, s, f): kDctTiles of gp_kernels.hpp, in order
SHAPES = ((64, 64), (128, 64), (64, 128), (128, 128), (256, 256), (512, 512), (1024, 1024))


def tile_time(nx: int, ny: int, precision: int, mode: int, tile: int, evaluations: int, reps: int):
    lib = hb.library()
    us = C.c_double()
    st = lib.mp_gp_dct_tile_time(nx, ny, precision, mode, tile, evaluations, reps, C.byref(us))
    if st == 1:
        return None  # the tile does not divide this product
    if st != 0:
        raise RuntimeError(lib.mp_gp_last_error().decode())
    return us.value


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=40)
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "dct_tiles.inc"))
    args = ap.parse_args()
    lib = hb.library()
    lib.mp_gp_dct_tile_time.argtypes = [C.c_int32] * 7 + [C.POINTER(C.c_double)]
    rows, summary = [], []
    for nx, ny in SHAPES:
        for precision in (0, 1):
            for mode in range(4):
                times = {}
                for tile in range(len(TILES)):
                    t1 = tile_time(nx, ny, precision, mode, tile, 1, args.reps)
                    if t1 is None:
                        continue
                    t2 = tile_time(nx, ny, precision, mode, tile, 2, args.reps)
                    times[tile] = (t1, t2)
                best = min(times, key=lambda k: sum(times[k]))
                rows.append((mode, precision, ny, nx, best))
                summary.append((nx, ny, precision, mode, best, times))
                print(f"{nx}x{ny} {'bf16x3' if precision else 'fp32'} {MODES[mode]}: tile {best} {TILES[best]} "
                      f"{times[best][0]:.2f}/{times[best][1]:.2f} us (tile 0: {times[0][0]:.2f}/{times[0][1]:.2f})",
                      file=sys.stderr, flush=True)
    stamp = datetime.datetime.now().strftime("%Y-%m-%d")
    with open(args.out, "w") as f:
        f.write(f"// {{mode, precision, rows, cols, tile}}: written by myplacement_amd/autotune.py on an MI355X, {stamp}\n")
        f.write("// (mode 0 fold-cols, 1 fold-rows, 2 mirror-cols, 3 mirror-rows; precision 0 FP32, 1 split-BF16)\n")
        for mode, precision, r, c, tile in rows:
            f.write(f"{{{mode}, {precision}, {r}, {c}, {tile}}},\n")
    # markdown summary on stdout
    print("| grid | precision | product | chosen tile (tr, tc, w, s, f) | us, 1 / 2 evaluations | tile 0 us | gain |")
    print("|---|---|---|---|---|---|---|")
    for nx, ny, precision, mode, best, times in summary:
        b, z = times[best], times[0]
        print(f"| {nx}x{ny} | {'bf16x3' if precision else 'fp32'} | {MODES[mode]} | {best} {TILES[best]} | "
              f"{b[0]:.2f} / {b[1]:.2f} | {z[0]:.2f} / {z[1]:.2f} | {sum(z) / sum(b):.2f}x |")
    return 0


if __name__ == "__main__":
    sys.exit(main())
