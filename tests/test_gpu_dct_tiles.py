"""Every tile shape of the half-length DCT products against a float64 evaluation of the same product.

The evaluation chain runs each product with the tile its compiled table names for the grid
(myplacement_amd/csrc/dct_tiles.inc, written by the autotuner); a new table may pick any tile.  So
every tile of gp_kernels.hpp (kDctTiles) is run here through mp_gp_dct_tile_run on random operands
of the chain's layouts and compared with numpy float64 products written from the layouts'
definitions below:

  n = transform length, h = n / 2, p = parity (0 even, 1 odd), tiles of TC output columns
  fold-cols    C[r][2c + p] = sum_t (A[r][t] +- A[r][n-1-t]) F[p][t][c]          A: ny x nx, F: [2][h][h]
  fold-rows    ev/od[r][c]  = sum_t F[p][r][t] (D[t][c] +- D[n-1-t][c]); C[2r + p][c] = chat * norm / den
               (den = (pi c / W)^2 + (pi l / H)^2, 0 where den is 0; W = nx, H = ny, norm = 1 / (4 nx ny))
  mirror-cols  ev/od[r][c]  = sum_t D[r][2t + p] F[p][t][c], c < n;  cosine half (c < h):
               C[r][c] = ev + od, C[r][n-1-c] = ev - od; sine half: C[r][n + c'] = ev + od,
               C[r][2n-1-c'] = od - ev (c' = c - h)                             D: ny x 2nx, F: [2][h][n]
  mirror-rows  ev/od[r][c]  = sum_t F[p][r][t] D[2t + p][c]; C[2 (r nx + c)] = ev + od,
               C[2 ((n-1-r) nx + c)] = sigma (ev - od), two products (sigma +1, -1)  D: ny x 2nx

Tolerances (normwise relative, written here): 5e-6 for the FP32 products (measured 1e-7 to 6e-7 from
64^2 to 1024^2), 5e-5 for split-BF16 (measured 2e-6 to 9e-6).  Tiles that split the contraction the
same way (the same s) sum every element in the same order, so their outputs must be bit-identical
(the prefetch setting f and the tile's rows, columns and workgroup width never change the sums).
"""

import ctypes as C
import sys

import numpy as np
import pytest

from myplacement_amd import backend as hb

pytestmark = pytest.mark.gpu

TILES = ((32, 32, 1, 8), (32, 32, 1, 4), (16, 32, 1, 4), (32, 32, 1, 2), (32, 32, 1, 1), (32, 32, 2, 1), (16, 32, 2, 1),
         (16, 16, 4, 1))
TILES = tuple(t + (f,) for f in (0, 1) for t in TILES)  # (tr, tc, w, s, f), kDctTiles order
MODES = ("fold-cols", "fold-rows", "mirror-cols", "mirror-rows")
TOL = {0: 5e-6, 1: 5e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def valid(tile, mode, nx, ny):
    tr, tc, w, s, f = TILES[tile]
    hr, hc = {0: (ny, nx // 2), 1: (ny // 2, nx), 2: (ny, nx), 3: (ny // 2, nx)}[mode]
    return hr % tr == 0 and hc % (w * tc) == 0


def reference(mode, nx, ny, data, fac):
    """float64 product of `mode` on the entry point's operand layout (see the module docstring)."""
    grid = ny * 2 * nx
    D = data.astype(np.float64)
    F = fac.astype(np.float64)
    out = np.zeros(data.size)
    if mode == 0:
        n, h = nx, nx // 2
        A = D[:ny * nx].reshape(ny, nx)
        Fp = F[:2 * h * h].reshape(2, h, h)
        ev = (A[:, :h] + A[:, ::-1][:, :h]) @ Fp[0]
        od = (A[:, :h] - A[:, ::-1][:, :h]) @ Fp[1]
        C0 = out[:ny * nx].reshape(ny, nx)
        C0[:, 0::2] = ev
        C0[:, 1::2] = od
    elif mode == 1:
        n, h = ny, ny // 2
        B = D[:ny * nx].reshape(ny, nx)
        Fp = F[:2 * h * h].reshape(2, h, h)
        ev = Fp[0] @ (B[:h] + B[::-1][:h])
        od = Fp[1] @ (B[:h] - B[::-1][:h])
        chat = np.empty((ny, nx))
        chat[0::2] = ev
        chat[1::2] = od
        k = np.arange(nx)[None, :]
        l = np.arange(ny)[:, None]
        den = (np.pi * k / nx) ** 2 + (np.pi * l / ny) ** 2
        phi = np.where(den > 1e-9, chat / (4.0 * nx * ny) / np.where(den > 1e-9, den, 1.0), 0.0)
        out[:ny * nx] = phi.ravel()
    elif mode == 2:
        n, h = nx, nx // 2
        A = D[:ny * nx].reshape(ny, nx)
        Fp = F[:2 * h * n].reshape(2, h, n)
        ev = A[:, 0::2] @ Fp[0]
        od = A[:, 1::2] @ Fp[1]
        C0 = out[:ny * 2 * nx].reshape(ny, 2 * nx)
        C0[:, :h] = (ev + od)[:, :h]
        C0[:, n - 1 - np.arange(h)] = (ev - od)[:, :h]
        C0[:, n:n + h] = (ev + od)[:, h:]
        C0[:, 2 * n - 1 - np.arange(h)] = (od - ev)[:, h:]
    else:
        n, h = ny, ny // 2
        Fp = F[:2 * h * h].reshape(2, h, h)
        for e, sigma in ((0, 1.0), (1, -1.0)):
            B = D[e * grid:(e + 1) * grid].reshape(ny, 2 * nx)[:, :nx]
            ev = Fp[0] @ B[0::2]
            od = Fp[1] @ B[1::2]
            Ce = out[e * grid:(e + 1) * grid].reshape(ny, 2 * nx)
            Cv = Ce.reshape(ny, nx, 2)[:, :, 0]
            Cv[:h] = ev + od
            Cv[n - 1 - np.arange(h)] = sigma * (ev - od)
    return out


@pytest.fixture(scope="module")
def lib():
    L = hb.library()
    L.mp_gp_dct_tile_run.argtypes = [C.c_int32] * 5 + [C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float),
                                                        C.c_int64, C.POINTER(C.c_float)]
    L.mp_gp_dct_tile_run.restype = C.c_int32
    L.mp_gp_dct_tile_for.restype = C.c_int32
    return L


def run(lib, nx, ny, precision, mode, tile, data, fac):
    out = np.zeros(data.size, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    st = lib.mp_gp_dct_tile_run(nx, ny, precision, mode, tile, data.ctypes.data_as(fp), data.size,
                                fac.ctypes.data_as(fp), fac.size, out.ctypes.data_as(fp))
    assert st == 0, hb.library().mp_gp_last_error().decode()
    return out


@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 64), (64, 128), (256, 256), (512, 512), (1024, 1024)])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp32", "bf16x3"])
def test_every_tile_matches_float64_products(lib, nx, ny, precision):
    rng = np.random.default_rng(nx * 7 + ny + precision)
    n = max(nx, ny)
    fac = rng.uniform(-1.0, 1.0, 2 * n * n).astype(np.float32)
    for mode in range(4):
        batch = 2 if mode == 3 else 1
        data = rng.uniform(-1.0, 1.0, batch * ny * 2 * nx).astype(np.float32)
        ref = reference(mode, nx, ny, data, fac)
        by_split = {}
        worst = 0.0
        tuned = lib.mp_gp_dct_tile_for(nx, ny, precision, mode)
        tiles = [t for t in range(len(TILES)) if valid(t, mode, nx, ny)]
        assert tuned in tiles
        for tile in tiles:
            say(f"[gpu] tile {tile} {TILES[tile]} {MODES[mode]} {nx}x{ny} precision {precision}")
            got = run(lib, nx, ny, precision, mode, tile, data, fac)
            err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            worst = max(worst, err)
            assert err < TOL[precision], (tile, MODES[mode], err)
            s = TILES[tile][3]
            if s in by_split:
                assert np.array_equal(got, by_split[s][1]), (tile, by_split[s][0], MODES[mode])
            else:
                by_split[s] = (tile, got)
        say(f"  {MODES[mode]}: {len(tiles)} tiles, worst normwise error {worst:.2e}, tuned tile {tuned}")


def test_tile_run_rejects_mismatched_operands(lib):
    data = np.zeros(64 * 2 * 64, dtype=np.float32)
    fac = np.zeros(2 * 64 * 64, dtype=np.float32)
    out = np.zeros(data.size, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    assert lib.mp_gp_dct_tile_run(64, 64, 0, 0, 0, data.ctypes.data_as(fp), data.size - 1, fac.ctypes.data_as(fp),
                                  fac.size, out.ctypes.data_as(fp)) != 0
    assert "sizes" in hb.library().mp_gp_last_error().decode()
    assert lib.mp_gp_dct_tile_run(64, 64, 0, 0, 99, data.ctypes.data_as(fp), data.size, fac.ctypes.data_as(fp),
                                  fac.size, out.ctypes.data_as(fp)) != 0
