"""The oracle restatement against the committed golden fixtures (exact
direct sums on seeded inputs)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "wstack_*.npz"))))
@pytest.mark.parametrize("eps", [1e-5, 1e-9])
def test_oracle_restatement_vs_golden(path, eps):
    g = np.load(path)
    nx, ny = g["dirty"].shape
    d = O.ms2dirty(g["uvw"], float(g["freq"]), g["vis"], g["wgt"], nx, ny, float(g["fov"]), eps)
    assert rel_l2(d, g["dft_dirty"]) < 5 * eps
    v = O.dirty2ms(g["uvw"], float(g["freq"]), g["dirty"], g["wgt"], float(g["fov"]), eps)
    assert rel_l2(v, g["dft_vis"]) < 100 * eps


@pytest.mark.parametrize("eps", [1e-5, 1e-9])
def test_oracle_nufft2d_vs_golden(eps):
    g = np.load(os.path.join(GOLDEN, "nufft2d_500_48x32.npz"))
    N2, N1 = g["fk"].shape
    assert rel_l2(O.nufft2d1(g["x"], g["y"], g["c"], N1, N2, eps), g["dft_type1"]) < 10 * eps
    assert rel_l2(O.nufft2d2(g["x"], g["y"], g["fk"], eps), g["dft_type2"]) < 10 * eps
