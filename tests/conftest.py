import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and libcurafft.so")


def rel_l2(a, b):
    a = np.ravel(np.asarray(a, dtype=np.complex128))
    b = np.ravel(np.asarray(b, dtype=np.complex128))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="session")
def gpu():
    """Skips when this machine has no GPU; on a GPU machine the library must
    load (a missing build is an error, not a skip)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU on this machine")
    import cugridder_amd
    cugridder_amd.lib()
    return torch


def as_float_abi(x):
    """A scalar or array as the float entry points receive it (FP32-rounded,
    then widened): the reference's float build plans kw from the float eps
    (float(1e-5) gives kw 7, not 6) and scales FP32 uvw."""
    if np.isscalar(x):
        return float(np.float32(x))
    a = np.asarray(x)
    if np.iscomplexobj(a):
        return a.astype(np.complex64).astype(np.complex128)
    return a.astype(np.float32).astype(np.float64)


def wst_case(nrow, nxdirty, nydirty, fov, seed=42, f0=1e9, use_wgt=True):
    """Inputs generated as the reference's python/curagridder/test_wst.py:55-66."""
    rng = np.random.default_rng(seed)
    c = 299792458.0
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / c)
    ms = rng.random(nrow) - 0.5 + 1j * (rng.random(nrow) - 0.5)
    dirty = rng.random((nxdirty, nydirty)) - 0.5
    wgt = rng.random(nrow) if use_wgt else None
    xpix = fov * np.pi / 180 / nxdirty
    ypix = fov * np.pi / 180 / nydirty
    return dict(uvw=uvw, ms=ms, dirty=dirty, wgt=wgt, f0=f0, xpix=xpix, ypix=ypix,
                fov=fov, nx=nxdirty, ny=nydirty)


def earth_rotation_rows(nrow, seed=3, nant=256, radius=4000.0, wmax_wl=5000.0, freq=1.4e9):
    """uvw rows (metres) of the full W-stacking workload's array: nant
    antennas uniform in a disk (latitude 45 deg), Earth-rotation tracks over
    8 h at 10 s steps toward dec 45 deg, both baseline orientations; nrow of
    the 65280 x 2880 samples drawn without replacement, then scaled so that
    max |w| = wmax_wl wavelengths at freq."""
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.random(nant))
    th = 2 * np.pi * rng.random(nant)
    E, N = r * np.cos(th), r * np.sin(th)
    lat = dec = np.deg2rad(45.0)
    X, Y, Z = -np.sin(lat) * N, E, np.cos(lat) * N
    ii, jj = np.triu_indices(nant, 1)
    bx = np.concatenate([X[jj] - X[ii], X[ii] - X[jj]])
    by = np.concatenate([Y[jj] - Y[ii], Y[ii] - Y[jj]])
    bz = np.concatenate([Z[jj] - Z[ii], Z[ii] - Z[jj]])
    nh = 2880
    sel = rng.choice(bx.size * nh, nrow, replace=False)
    b, t = sel // nh, sel % nh
    H = np.linspace(-np.pi / 3, np.pi / 3, nh)[t]
    sH, cH = np.sin(H), np.cos(H)
    u = sH * bx[b] + cH * by[b]
    v = -np.sin(dec) * cH * bx[b] + np.sin(dec) * sH * by[b] + np.cos(dec) * bz[b]
    w = np.cos(dec) * cH * bx[b] - np.cos(dec) * sH * by[b] + np.sin(dec) * bz[b]
    uvw = np.stack([u, v, w], axis=1)
    uvw *= wmax_wl * (299792458.0 / freq) / np.abs(uvw[:, 2]).max()
    return uvw
