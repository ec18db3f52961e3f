"""Writes the FP64 golden vectors of the Neumann field path with scipy.fft.

scipy.fft.dct(type=2), dct(type=3) and dst(type=3) with norm=None are FFTW's REDFT10, REDFT01 and
RODFT01 by definition (unnormalised), the three transforms of ElectrostaticField.cpp:164-181.  The
field solve below restates ElectrostaticField.cpp:205-237 on top of them, independently of the
oracle's own transforms.  Run from the repository root:  python tests/golden/make_golden.py
"""

import os

import numpy as np
from scipy import fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def neumann(source, region):
    rows, cols = source.shape
    W, H = region[2] - region[0], region[3] - region[1]
    src = source - source.mean()
    coeff = sf.dct(sf.dct(src, type=2, axis=1), type=2, axis=0)
    wx = np.pi * np.arange(cols) / W
    wy = np.pi * np.arange(rows) / H
    den = wy[:, None] ** 2 + wx[None, :] ** 2
    norm = 1.0 / (4.0 * cols * rows)
    pc = np.where(den > 1e-9, coeff * norm / np.where(den > 1e-9, den, 1.0), 0.0)
    fxc = np.zeros_like(pc)
    fyc = np.zeros_like(pc)
    fxc[:, :-1] = (pc * wx[None, :])[:, 1:]
    fyc[:-1, :] = (pc * wy[:, None])[1:, :]
    pot = sf.dct(sf.dct(pc, type=3, axis=1), type=3, axis=0)
    fx = sf.dst(sf.dct(fxc, type=3, axis=0), type=3, axis=1)
    fy = sf.dct(sf.dst(fyc, type=3, axis=0), type=3, axis=1)
    return src, pot, fx, fy


def main():
    rng = np.random.default_rng(20261203)
    out = {}
    for n in (12, 16, 64, 100):
        x = rng.standard_normal(n)
        out[f"x{n}"] = x
        out[f"redft10_{n}"] = sf.dct(x, type=2)
        out[f"redft01_{n}"] = sf.dct(x, type=3)
        out[f"rodft01_{n}"] = sf.dst(x, type=3)
    for rows, cols, region in ((12, 16, (-3.0, 2.0, 5.0, 8.0)), (64, 64, (0.0, 0.0, 400.0, 360.0)),
                               (48, 80, (10.0, -5.0, 1210.0, 10795.0))):
        s = rng.random((rows, cols)) * 1.5 - 0.4
        src, pot, fx, fy = neumann(s, region)
        key = f"field_{rows}x{cols}"
        out[key + "_in"] = s
        out[key + "_region"] = np.array(region)
        out[key + "_potential"] = pot
        out[key + "_fx"] = fx
        out[key + "_fy"] = fy
    np.savez_compressed(os.path.join(HERE, "neumann_golden.npz"), **out)


if __name__ == "__main__":
    main()
