"""Writes the golden fixtures in this directory.

The reference ships no golden vectors (its tests print DFT errors and assert
only adjointness), so the fixtures hold seeded inputs and the exact answer of
the mathematics: the FP64 direct sum (python/curagridder/test_wst.py:15-32
formula, C-ABI sign convention; the 2-D NUFFT sums of
test/nufft_2d_1_test.cu:276-283 and test/nufft_2d_2_test.cu:316-326).
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def wstack_case(name, nrow, nx, ny, fov, f0, seed, gaussian=False):
    rng = np.random.default_rng(seed)
    c = 299792458.0
    if gaussian:
        uvw = np.empty((nrow, 3))
        uvw[:, 0] = rng.normal(0, 1500, nrow)
        uvw[:, 1] = rng.normal(0, 1500, nrow)
        uvw[:, 2] = rng.uniform(-300, 300, nrow)
    else:
        uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / c)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    wgt = rng.uniform(0.5, 1.5, nrow)
    dirty = rng.standard_normal((nx, ny)) + 0j
    xpix = fov * np.pi / 180 / nx
    ypix = fov * np.pi / 180 / ny
    dft_dirty = O.dft_ms2dirty(uvw, f0, vis, wgt, nx, ny, xpix, ypix, sign=-1).reshape(nx, ny)
    dft_vis = O.dft_dirty2ms(uvw, f0, dirty, wgt, xpix, ypix, sign=-1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), uvw=uvw, vis=vis, wgt=wgt, dirty=dirty,
                        freq=f0, fov=fov, dft_dirty=dft_dirty, dft_vis=dft_vis)


def nufft_case(name, M, N1, N2, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-np.pi, np.pi, M)
    y = rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fk = rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), x=x, y=y, c=c, fk=fk,
                        dft_type1=O.dft2d1(x, y, c, N1, N2).reshape(N2, N1),
                        dft_type2=O.dft2d2(x, y, fk))


if __name__ == "__main__":
    wstack_case("wstack_uniform_300_32x64_fov10", 300, 32, 64, 10.0, 1e9, 11)
    wstack_case("wstack_gauss_2000_128x96_fov2", 2000, 128, 96, 2.0, 1e8, 12, gaussian=True)
    wstack_case("wstack_uniform_50_250x64_fov20", 50, 250, 64, 20.0, 1e9, 13)
    nufft_case("nufft2d_500_48x32", 500, 48, 32, 14)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
