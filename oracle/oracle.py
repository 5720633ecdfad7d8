"""ctypes wrapper of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference's ms2dirty/dirty2ms algorithm in FP64 C
(``cugrid_oracle.c``) and provides an FP64 direct-DFT for both directions.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module; the product library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
C_LIGHT = 299792458.0

_dp = np.ctypeslib.ndpointer(np.float64, flags="C")
_cp = np.ctypeslib.ndpointer(np.complex128, flags="C")
_ip = np.ctypeslib.ndpointer(np.int64, flags="C")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.orc_next235beven.argtypes = [c_int, c_int]
        L.orc_next235beven.restype = c_int
        L.orc_gauss_legendre.argtypes = [c_int, _dp, _dp]
        L.orc_setup_conv.argtypes = [c_double, c_double, c_int,
                                     ctypes.POINTER(c_int),
                                     ctypes.POINTER(c_double),
                                     ctypes.POINTER(c_double)]
        L.orc_setup_conv.restype = c_int
        L.orc_num_cells.argtypes = [c_int, c_double, c_int]
        L.orc_num_cells.restype = c_int
        L.orc_set_nhg_w.argtypes = [c_double, c_double, c_int, c_double,
                                    ctypes.POINTER(c_int),
                                    ctypes.POINTER(c_double),
                                    ctypes.POINTER(c_double)]
        L.orc_fseries_grid.argtypes = [c_int, c_double, c_double, c_int, _dp]
        L.orc_fseries_at.argtypes = [c_int, c_double, c_double, _dp, c_int, _dp]
        L.orc_fft1d.argtypes = [c_int, _cp, c_int]
        L.orc_fft2d.argtypes = [c_int, c_int, c_int, _cp, c_int]
        L.orc_spread3d.argtypes = [c_int, _dp, _dp, _dp, _cp, c_int, c_int,
                                   c_int, c_int, c_double, c_double, c_int, _cp]
        L.orc_interp3d.argtypes = [c_int, _dp, _dp, _dp, _cp, c_int, c_int,
                                   c_int, c_int, c_double, c_double, c_int, _cp]
        L.orc_ms2dirty.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                   _dp, _cp, ctypes.c_void_p, _cp, c_double,
                                   c_double, c_int, c_int]
        L.orc_ms2dirty.restype = c_int
        L.orc_dirty2ms.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                   _dp, _cp, ctypes.c_void_p, _cp, c_double,
                                   c_double, c_int, c_int]
        L.orc_dirty2ms.restype = c_int
        L.orc_dft_ms2dirty.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                       c_double, _dp, _cp, ctypes.c_void_p,
                                       c_int, c_int, _ip, _cp]
        L.orc_dft_dirty2ms.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                       c_double, _dp, _cp, ctypes.c_void_p,
                                       c_int, c_int, _ip, _cp]
        L.orc_dft_dirty2ms_sparse.argtypes = [c_int, c_int, c_double, c_double, c_double, _dp,
                                              c_int, _ip, _cp, ctypes.c_void_p, c_int, c_int,
                                              _ip, _cp]
        L.orc_nufft2d1.argtypes = [c_int, _dp, _dp, _cp, c_int, c_int,
                                   c_double, c_double, c_int, _cp]
        L.orc_nufft2d1.restype = c_int
        L.orc_nufft2d2.argtypes = [c_int, _dp, _dp, _cp, c_int, c_int,
                                   c_double, c_double, c_int, _cp]
        L.orc_nufft2d2.restype = c_int
        L.orc_dft2d1.argtypes = [c_int, _dp, _dp, _cp, c_int, c_int, c_int,
                                 _ip, _cp]
        L.orc_dft2d2.argtypes = [_dp, _dp, c_int, c_int, _cp, c_int, _ip, _cp]
        L.orc_plan_export.argtypes = [c_int, c_int, c_int, c_double, c_double,
                                      _dp, c_double, c_double, c_int, c_int, _dp]
        L.orc_plan_export.restype = c_int
        L.orc_num_threads.restype = c_int
        L.orc_set_pix_y.argtypes = [c_double]
        L.orc_ms2dirty_grid.argtypes = [c_int, c_int, c_int, c_double, c_double, _dp, _cp,
                                        ctypes.c_void_p, c_double, c_double, c_int, c_int, _cp]
        L.orc_ms2dirty_grid.restype = c_int
        L.orc_dirty2ms_planes.argtypes = [c_int, c_int, c_int, c_double, c_double, _dp, _cp,
                                          c_double, c_double, c_int, c_int, _cp]
        L.orc_dirty2ms_planes.restype = c_int
        _LIB = L
    return _LIB


def _wptr(wgt):
    if wgt is None:
        return None, None
    w = np.ascontiguousarray(wgt, dtype=np.float64).ravel()
    return w, w.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------- planner

def next235beven(n, b=1):
    return lib().orc_next235beven(int(n), int(b))


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, x, w)
    return x, w


def setup_conv(eps, sigma=2.0, single=False):
    kw = ctypes.c_int()
    beta = ctypes.c_double()
    c = ctypes.c_double()
    ier = lib().orc_setup_conv(float(eps), float(sigma), int(single),
                               ctypes.byref(kw), ctypes.byref(beta),
                               ctypes.byref(c))
    return ier, kw.value, beta.value, c.value


def num_cells(n, sigma, kw):
    return lib().orc_num_cells(int(n), float(sigma), int(kw))


def set_nhg_w(S, X, kw, sigma=2.0):
    nf = ctypes.c_int()
    h = ctypes.c_double()
    g = ctypes.c_double()
    lib().orc_set_nhg_w(float(S), float(X), int(kw), float(sigma),
                        ctypes.byref(nf), ctypes.byref(h), ctypes.byref(g))
    return nf.value, h.value, g.value


PLAN_KEYS = ("nf1", "nf2", "nf3", "kw", "beta", "es_c", "pix_x", "pix_y",
             "i_center", "i_half_width", "o_center", "o_half_width", "gamma",
             "h", "f_over_c", "ier")


def plan(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma=2.0, sign=-1,
         single=False):
    out = np.zeros(16)
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    lib().orc_plan_export(nrow, nxdirty, nydirty, fov, freq, uvw, eps, sigma,
                          sign, int(single), out)
    d = dict(zip(PLAN_KEYS, out.tolist()))
    for k in ("nf1", "nf2", "nf3", "kw", "ier"):
        d[k] = int(d[k])
    return d


def fseries_grid(kw, beta, c, nf):
    out = np.zeros(nf // 2 + 1)
    lib().orc_fseries_grid(kw, beta, c, nf, out)
    return out


def fseries_at(kw, beta, c, k):
    k = np.ascontiguousarray(k, dtype=np.float64)
    out = np.zeros(k.size)
    lib().orc_fseries_at(kw, beta, c, k, k.size, out)
    return out


# --------------------------------------------------------------------- FFT

def fft1d(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(x.size, x, int(sign))
    return x


def fft2d(planes, sign):
    a = np.ascontiguousarray(planes, dtype=np.complex128).copy()
    if a.ndim == 2:
        a = a[None]
    lib().orc_fft2d(a.shape[0], a.shape[1], a.shape[2], a, int(sign))
    return a


def spread3d(x, y, z, c, nf1, nf2, nf3, kw, beta, es_c, pirange=0):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    c = np.ascontiguousarray(c, dtype=np.complex128)
    fw = np.zeros((nf3, nf2, nf1), dtype=np.complex128)
    lib().orc_spread3d(x.size, x, y, z, c, nf1, nf2, nf3, kw, beta, es_c,
                       pirange, fw)
    return fw


def interp3d(x, y, z, fw, kw, beta, es_c, pirange=0):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    fw = np.ascontiguousarray(fw, dtype=np.complex128)
    nf3, nf2, nf1 = fw.shape
    c = np.zeros(x.size, dtype=np.complex128)
    lib().orc_interp3d(x.size, x, y, z, c, nf1, nf2, nf3, kw, beta, es_c,
                       pirange, fw)
    return c


# ---------------------------------------------------- W-stacking pipelines

def ms2dirty(uvw, freq, vis, wgt, nxdirty, nydirty, fov, eps, sigma=2.0,
             sign=-1, single=False, pix_y=None):
    """FP64 restatement of ms2dirty_*; returns complex dirty [nx, ny].
    pix_y (radians) replaces the reference's fov / ny pixel size along ny."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128).ravel()
    w, wp = _wptr(wgt)
    dirty = np.zeros((nxdirty, nydirty), dtype=np.complex128)
    lib().orc_set_pix_y(float(pix_y or 0.0))
    try:
        ier = lib().orc_ms2dirty(uvw.shape[0], nxdirty, nydirty, fov, freq, uvw,
                                 vis, wp, dirty, eps, sigma, sign, int(single))
    finally:
        lib().orc_set_pix_y(0.0)
    if ier >= 2:
        raise ValueError("oracle ms2dirty: bad parameters (ier=%d)" % ier)
    return dirty


def dirty2ms(uvw, freq, dirty, wgt, fov, eps, sigma=2.0, sign=-1,
             single=False, pix_y=None):
    """FP64 restatement of dirty2ms_*; returns complex vis [nrow]."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    dirty = np.ascontiguousarray(dirty, dtype=np.complex128)
    nx, ny = dirty.shape
    w, wp = _wptr(wgt)
    vis = np.zeros(uvw.shape[0], dtype=np.complex128)
    lib().orc_set_pix_y(float(pix_y or 0.0))
    try:
        ier = lib().orc_dirty2ms(uvw.shape[0], nx, ny, fov, freq, uvw, vis, wp,
                                 dirty, eps, sigma, sign, int(single))
    finally:
        lib().orc_set_pix_y(0.0)
    if ier >= 2:
        raise ValueError("oracle dirty2ms: bad parameters (ier=%d)" % ier)
    return vis


def _wrap_pi(x):
    """x mod 2 pi in [-pi, pi) (SHIFT_RESCALE's wrap, include/conv.h:24)."""
    return x - 2 * np.pi * np.floor((x + np.pi) / (2 * np.pi))


def uv_coords_2d(uvw, freq, pix_x, pix_y):
    """Plain 2-D gridding coordinates in [-pi, pi): x (grid axis 1, image
    columns, ny) = 2 pi f/c v pix_y and y (axis 2, rows, nx) = 2 pi f/c u
    pix_x, as the W-stacking coordinates of cugridder.cu:191,201-202
    without w."""
    uvw = np.asarray(uvw, dtype=np.float64)
    s = 2 * np.pi * freq / C_LIGHT
    return _wrap_pi(uvw[:, 1] * s * pix_y), _wrap_pi(uvw[:, 0] * s * pix_x)


def ms2dirty_2d(uvw, freq, vis, wgt, nxdirty, nydirty, pix_x, pix_y, eps, sigma=2.0,
                single=False):
    """Plain 2-D gridding (NIFTY do_wstacking=False: w ignored, no 1/n):
    dirty(l, m) = sum_j wgt_j vis_j e^{2 pi i f/c (u l + v m)}, the 2-D
    NUFFT type 1 restatement (conv_2d_nputsdriven + deconv_2d) on the
    scaled (v, u); returns complex dirty [nx, ny]."""
    x, y = uv_coords_2d(uvw, freq, pix_x, pix_y)
    c = np.asarray(vis, dtype=np.complex128).ravel()
    if wgt is not None:
        c = c * np.asarray(wgt, dtype=np.float64).ravel()
    return nufft2d1(x, y, c, nydirty, nxdirty, eps, sigma, single)


def dirty2ms_2d(uvw, freq, dirty, wgt, pix_x, pix_y, eps, sigma=2.0, single=False):
    """Adjoint of ms2dirty_2d: vis_j = wgt_j sum_{l,m} dirty e^{-2 pi i f/c
    (u l + v m)}, the 2-D NUFFT type 2 restatement."""
    x, y = uv_coords_2d(uvw, freq, pix_x, pix_y)
    v = nufft2d2(x, y, np.asarray(dirty, dtype=np.complex128), eps, sigma, single)
    if wgt is not None:
        v = v * np.asarray(wgt, dtype=np.float64).ravel()
    return v


def ms2dirty_grid(uvw, freq, vis, wgt, nxdirty, nydirty, fov, eps, sigma=2.0, sign=-1,
                  single=False):
    """ms2dirty's w-plane grid after spreading, before the FFTs [nf3, nf2, nf1]."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    p = plan(uvw.shape[0], nxdirty, nydirty, fov, freq, uvw, eps, sigma, sign, single)
    vis = np.ascontiguousarray(vis, dtype=np.complex128).ravel()
    w, wp = _wptr(wgt)
    fw = np.zeros((p["nf3"], p["nf2"], p["nf1"]), dtype=np.complex128)
    lib().orc_ms2dirty_grid(uvw.shape[0], nxdirty, nydirty, fov, freq, uvw, vis, wp, eps, sigma,
                            sign, int(single), fw)
    return fw


def dirty2ms_planes(uvw, freq, dirty, fov, eps, sigma=2.0, sign=-1, single=False):
    """dirty2ms's w-plane grid after the w-IDFT, before the FFTs."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    dirty = np.ascontiguousarray(dirty, dtype=np.complex128)
    nx, ny = dirty.shape
    p = plan(uvw.shape[0], nx, ny, fov, freq, uvw, eps, sigma, sign, single)
    fw = np.zeros((p["nf3"], p["nf2"], p["nf1"]), dtype=np.complex128)
    lib().orc_dirty2ms_planes(uvw.shape[0], nx, ny, fov, freq, uvw, dirty, eps, sigma, sign,
                              int(single), fw)
    return fw


def dft_ms2dirty(uvw, freq, vis, wgt, nxdirty, nydirty, pix_x, pix_y, sign=-1,
                 pixels=None):
    """Direct FP64 DFT image (complex), optionally on a pixel subset."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    vis = np.ascontiguousarray(vis, dtype=np.complex128).ravel()
    w, wp = _wptr(wgt)
    if pixels is None:
        pixels = np.arange(nxdirty * nydirty, dtype=np.int64)
    pixels = np.ascontiguousarray(pixels, dtype=np.int64)
    out = np.zeros(pixels.size, dtype=np.complex128)
    lib().orc_dft_ms2dirty(uvw.shape[0], nxdirty, nydirty, pix_x, pix_y, freq,
                           uvw, vis, wp, sign, pixels.size, pixels, out)
    return out


def dft_dirty2ms(uvw, freq, dirty, wgt, pix_x, pix_y, sign=-1, rows=None):
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    dirty = np.ascontiguousarray(dirty, dtype=np.complex128)
    nx, ny = dirty.shape
    w, wp = _wptr(wgt)
    if rows is None:
        rows = np.arange(uvw.shape[0], dtype=np.int64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros(rows.size, dtype=np.complex128)
    lib().orc_dft_dirty2ms(uvw.shape[0], nx, ny, pix_x, pix_y, freq, uvw,
                           dirty, wp, sign, rows.size, rows, out)
    return out


def dft_dirty2ms_sparse(uvw, freq, nxdirty, nydirty, pixels, values, wgt, pix_x, pix_y, sign=-1,
                        rows=None):
    """dft_dirty2ms of an image that is zero except values at the flat pixel
    indices `pixels` (row * nydirty + col)."""
    uvw = np.ascontiguousarray(uvw, dtype=np.float64)
    pixels = np.ascontiguousarray(pixels, dtype=np.int64)
    values = np.ascontiguousarray(values, dtype=np.complex128)
    w, wp = _wptr(wgt)
    if rows is None:
        rows = np.arange(uvw.shape[0], dtype=np.int64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros(rows.size, dtype=np.complex128)
    lib().orc_dft_dirty2ms_sparse(nxdirty, nydirty, pix_x, pix_y, freq, uvw, pixels.size, pixels,
                                  values, wp, sign, rows.size, rows, out)
    return out


# ------------------------------------------------------------- 2-D NUFFT

def nufft2d1(x, y, c, N1, N2, eps, sigma=2.0, single=False):
    x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
    c = np.ascontiguousarray(c, dtype=np.complex128)
    fk = np.zeros((N2, N1), dtype=np.complex128)
    lib().orc_nufft2d1(x.size, x, y, c, N1, N2, eps, sigma, int(single), fk)
    return fk


def nufft2d2(x, y, fk, eps, sigma=2.0, single=False):
    x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
    fk = np.ascontiguousarray(fk, dtype=np.complex128)
    N2, N1 = fk.shape
    c = np.zeros(x.size, dtype=np.complex128)
    lib().orc_nufft2d2(x.size, x, y, c, N1, N2, eps, sigma, int(single), fk)
    return c


def dft2d1(x, y, c, N1, N2, modes=None):
    x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
    c = np.ascontiguousarray(c, dtype=np.complex128)
    if modes is None:
        modes = np.arange(N1 * N2, dtype=np.int64)
    modes = np.ascontiguousarray(modes, dtype=np.int64)
    out = np.zeros(modes.size, dtype=np.complex128)
    lib().orc_dft2d1(x.size, x, y, c, N1, N2, modes.size, modes, out)
    return out


def dft2d2(x, y, fk, pts=None):
    x, y = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y))
    fk = np.ascontiguousarray(fk, dtype=np.complex128)
    N2, N1 = fk.shape
    if pts is None:
        pts = np.arange(x.size, dtype=np.int64)
    pts = np.ascontiguousarray(pts, dtype=np.int64)
    out = np.zeros(pts.size, dtype=np.complex128)
    lib().orc_dft2d2(x, y, N1, N2, fk, pts.size, pts, out)
    return out


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------- reference GLR (oracle/_ref)

def ref_glr(n):
    """Gauss-Legendre rule from the reference's own contrib/legendre_rule_fast.c
    compiled into oracle/_ref (None when the reference tree was absent)."""
    path = os.path.join(_HERE, "_ref", "libglr_ref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.legendre_compute_glr.argtypes = [ctypes.c_int, _dp, _dp]
    x = np.zeros(n)
    w = np.zeros(n)
    L.legendre_compute_glr(n, x, w)
    return x, w
