"""CPU checks of the drop-in boundary: libcurafft.so loads, exports every
function include/curafft.h declares, and its host-only planner agrees with
the oracle's restatement of the reference planner."""
import ctypes
import os
import re

import numpy as np
import pytest

import cugridder_amd
from oracle import oracle as O
from tests.conftest import ROOT, wst_case

HEADER = os.path.join(ROOT, "include", "curafft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_reference_symbols():
    names = declared_functions()
    for ref in ("ms2dirty_1", "ms2dirty_2", "dirty2ms_1", "dirty2ms_2", "ms2dirtyf_1",
                "ms2dirtyf_2", "dirty2msf_1", "dirty2msf_2"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(cugridder_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_argtypes():
    L = cugridder_amd.lib()
    assert L.cg_version() == 102
    for n in declared_functions():
        assert getattr(L, n).restype is ctypes.c_int


@pytest.mark.parametrize("n", [1, 2, 96, 128, 160, 200, 320, 1250, 2000, 4096, 6250, 13122, 16384])
def test_fft_col_split(n):
    """Columns longer than 128 are split n = na * nb (both <= 256); the
    spectrum row map is then a permutation with frequency ka + na kb at row
    ka nb + kb."""
    from cugridder_amd import Context
    L = cugridder_amd.lib()
    na, nb = ctypes.c_int(), ctypes.c_int()
    r = L.cg_fft_col_split(n, ctypes.byref(na), ctypes.byref(nb))
    rows = Context.spectrum_rows(n)
    assert sorted(rows.tolist()) == list(range(n))
    if n <= 128:
        assert r == 0 and np.array_equal(rows, np.arange(n))
        return
    assert r == 1
    a, b = na.value, nb.value
    assert a * b == n and 2 <= a <= 256 and 2 <= b <= 256
    for ka, kb in ((0, 0), (a - 1, 0), (0, b - 1), (a - 1, b - 1), (a // 2, b // 3)):
        assert rows[ka * b + kb] == ka + a * kb
    assert L.cg_fft_col_split(0, ctypes.byref(na), ctypes.byref(nb)) == 2


@pytest.mark.parametrize("nrow,nx,ny,fov,eps,freq", [
    (100, 32, 64, 10.0, 1e-10, 1e9), (27, 16, 250, 20.0, 5e-7, 1e9), (2, 64, 128, 1.0, 5e-3, 1e9),
    (1000, 1024, 1024, 2.0, 1e-5, 1e8), (500, 8192, 8192, 3.0, 1e-6, 1.4e9), (10, 250, 64, 1.0, 2e-1, 1e9)])
@pytest.mark.parametrize("single", [False, True])
def test_planner_matches_oracle(nrow, nx, ny, fov, eps, freq, single):
    k = wst_case(nrow, nx, ny, fov, f0=freq)
    uvw = k["uvw"]
    ref = O.plan(nrow, nx, ny, fov, freq, uvw, eps, sign=-1, single=single)
    w = uvw[:, 2] * freq / 299792458.0 * -1 * 2 * np.pi
    out = np.zeros(16)
    ier = cugridder_amd.lib().cg_plan_query(nrow, nx, ny, fov, freq, w.min(), w.max(), eps, 2.0, -1,
                                           int(single), out)
    assert ier == ref["ier"]
    keys = ("nf1", "nf2", "nf3", "kw", "beta", "es_c", "pix_x", "pix_y", "i_center",
            "i_half_width", "o_center", "o_half_width", "gamma", "h", "f_over_c")
    for i, key in enumerate(keys):
        # i_center = (wmax + wmin)/2 cancels; both sides scale w in a
        # different multiplication order
        tol_abs = 1e-13 * ref["i_half_width"] if key == "i_center" else 1e-300
        assert out[i] == pytest.approx(ref[key], rel=1e-12, abs=tol_abs), key


def test_planner_rejects_bad_arguments():
    out = np.zeros(16)
    L = cugridder_amd.lib()
    assert L.cg_plan_query(10, 0, 64, 1.0, 1e9, 0.0, 1.0, 1e-5, 2.0, -1, 0, out) == 2
    assert L.cg_plan_query(10, 64, 64, -1.0, 1e9, 0.0, 1.0, 1e-5, 2.0, -1, 0, out) == 2
    assert L.cg_plan_query(10, 64, 64, 1.0, 1e9, 0.0, 1.0, 1e-5, 1.0, -1, 0, out) == 2
    # eps below the precision floor is clamped with ier = 1 (setup_conv_opts)
    assert L.cg_plan_query(10, 64, 64, 1.0, 1e9, 0.0, 1.0, 1e-9, 2.0, -1, 1, out) == 1


def test_nf3_known_values():
    """The w-plane count at the two W-stacking workloads, from the closed
    form nf3 = max(even(floor(2 sigma S X / pi) + kw + 1), 2 kw) worked by
    hand: 1024^2 at 2 deg, 1e8 Hz, |w| <= 300 m (S = 1.523e-4, X = 628.9,
    so 7.12 -> 8 < 2 kw = 12): 12 planes; 8192^2 at 3 deg, 1.4 GHz, |w| =
    5000 wavelengths (S = 3.428e-4, X = 2 pi 5000, 13.71 + 8 = 21.71): 22
    planes, and 18 / 24 at 4000 / 5500 wavelengths.  An over- or under-sized
    plane count fails here."""
    c = 299792458.0
    L = cugridder_amd.lib()
    out = np.zeros(16)
    w = 300.0 * 2 * np.pi * 1e8 / c
    assert L.cg_plan_query(131072, 1024, 1024, 2.0, 1e8, -w, w, 1e-5, 2.0, -1, 1, out) == 0
    assert int(out[2]) == 12 and int(out[3]) == 6
    for wl, nf3 in ((5000.0, 22), (4000.0, 18), (5500.0, 24)):
        w = 2 * np.pi * wl
        assert L.cg_plan_query(1 << 27, 8192, 8192, 3.0, 1.4e9, -w, w, 1e-6, 2.0, -1, 1, out) == 0
        assert (int(out[0]), int(out[2]), int(out[3])) == (16384, nf3, 7), wl
        uvw = np.array([[0.0, 0.0, wl * c / 1.4e9], [0.0, 0.0, -wl * c / 1.4e9]])
        assert O.plan(2, 8192, 8192, 3.0, 1.4e9, uvw, 1e-6)["nf3"] == nf3
