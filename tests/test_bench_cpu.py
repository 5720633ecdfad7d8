"""CPU checks of bench.py's measurement code: the FP64 direct sum it uses
to report accuracy agrees with the oracle's direct sum (both directions,
with and without weights), and the HBM model adds the kernel table's bytes
over launches."""
import numpy as np
import pytest
import torch

import bench
from oracle import oracle as O
from tests.conftest import rel_l2, wst_case


@pytest.mark.parametrize("use_wgt", [False, True])
def test_bench_direct_sum_matches_oracle(use_wgt):
    nx, ny, fov = 24, 32, 6.0
    k = wst_case(300, nx, ny, fov, use_wgt=use_wgt)
    vis = k["ms"].astype(np.complex64)
    img = (k["dirty"] + 0.5j * k["dirty"][::-1]).astype(np.complex64)
    W = dict(uvw=torch.from_numpy(k["uvw"]), vis=torch.from_numpy(vis),
             wgt=None if k["wgt"] is None else torch.from_numpy(k["wgt"].astype(np.float32)),
             img=torch.from_numpy(img), nx=nx, ny=ny, fov=fov, freq=k["f0"])
    w64 = None if k["wgt"] is None else k["wgt"].astype(np.float32).astype(np.float64)
    pix = fov * np.pi / 180
    ref_d = O.dft_ms2dirty(k["uvw"], k["f0"], vis.astype(np.complex128), w64, nx, ny, pix / nx,
                           pix / ny).reshape(nx, ny)
    ref_v = O.dft_dirty2ms(k["uvw"], k["f0"], img.astype(np.complex128), w64, pix / nx, pix / ny)
    # exact outputs give zero error; a perturbed output gives its perturbation
    acc = bench.dft_accuracy(torch, W, torch.from_numpy(ref_d), torch.from_numpy(ref_v),
                             npix_sample=nx * ny, nrow_sample=300)
    assert acc["ms2dirty_rel_l2_vs_dft"] < 1e-12 and acc["dirty2ms_rel_l2_vs_dft"] < 1e-12
    acc = bench.dft_accuracy(torch, W, torch.from_numpy(ref_d * (1 + 1e-3)),
                             torch.from_numpy(ref_v * (1 - 2e-3)), npix_sample=nx * ny,
                             nrow_sample=300)
    assert acc["ms2dirty_rel_l2_vs_dft"] == pytest.approx(1e-3, rel=1e-6)
    assert acc["dirty2ms_rel_l2_vs_dft"] == pytest.approx(2e-3, rel=1e-6)


def test_hbm_model_sums_launches():
    W = {"M": 1000}
    kernels = [("a", "ms2dirty", 1, 1.0, 4e9), ("b", "ms2dirty", 2, 1.0, 2e9)]
    m = bench.hbm_model(W, kernels, ms_per_step=2.0)
    assert m["step_algorithmic_GB"] == 8.0
    assert m["step_frac_of_hbm_peak"] == pytest.approx(8e9 / 2e-3 / 8e12)
    assert m["ceiling_Mvis_s_at_peak"] == pytest.approx(1000 / (8e9 / 8e12) / 1e6, rel=1e-3)
