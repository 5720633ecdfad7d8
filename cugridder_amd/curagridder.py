"""Plain complex-in/complex-out vis2dirty / dirty2vis over libcurafft.so.

Mirrors the reference binding python/curagridder/curagridder.py:92-151
(sign = +1, complex dirty image passed in and returned, fov in degrees), plus
the mixed-precision and 2-D NUFFT entry points of this library.  The
reference declares ms2dirtyf_2's fov/freq as c_double although the C ABI is
float (curagridder.py:77); here the float ABI is declared as float.
"""
import numpy as np

from ._lib import check, lib

__all__ = ["vis2dirty", "dirty2vis", "vis2dirty_mixed", "dirty2vis_mixed", "nufft2d1",
           "nufft2d2"]


def vis2dirty(uvw, freq, ms, wgt, dirty, fov, epsilon=1e-6, sigma=2):
    """dirty (complex, (nx, ny)) <- visibilities; returns dirty."""
    L = lib()
    nrow = uvw.shape[0]
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sign = 1
    f0 = float(np.asarray(freq).ravel()[0])
    if dirty.dtype == np.complex128:
        args = (nrow, nxdirty, nydirty, fov, f0, np.ascontiguousarray(uvw, np.float64),
                np.ascontiguousarray(ms, np.complex128))
        if wgt is None:
            ier = L.ms2dirty_1(*args, dirty, epsilon, sigma, sign)
        else:
            ier = L.ms2dirty_2(*args, np.ascontiguousarray(wgt, np.float64), dirty, epsilon,
                               sigma, sign)
    else:
        args = (nrow, nxdirty, nydirty, fov, f0, np.ascontiguousarray(uvw, np.float32),
                np.ascontiguousarray(ms, np.complex64))
        if wgt is None:
            ier = L.ms2dirtyf_1(*args, dirty, epsilon, sigma, sign)
        else:
            ier = L.ms2dirtyf_2(*args, np.ascontiguousarray(wgt, np.float32), dirty, epsilon,
                                sigma, sign)
    check(ier, "vis2dirty")
    return dirty


def dirty2vis(uvw, freq, ms, wgt, dirty, fov, epsilon=1e-6, sigma=2):
    """visibilities ms (complex, written in place) <- dirty; returns ms."""
    L = lib()
    nrow = uvw.shape[0]
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sign = 1
    f0 = float(np.asarray(freq).ravel()[0])
    if ms.dtype == np.complex128:
        args = (nrow, nxdirty, nydirty, fov, f0, np.ascontiguousarray(uvw, np.float64), ms)
        d = np.ascontiguousarray(dirty, np.complex128)
        if wgt is None:
            ier = L.dirty2ms_1(*args, d, epsilon, sigma, sign)
        else:
            ier = L.dirty2ms_2(*args, np.ascontiguousarray(wgt, np.float64), d, epsilon, sigma,
                               sign)
    else:
        args = (nrow, nxdirty, nydirty, fov, f0, np.ascontiguousarray(uvw, np.float32), ms)
        d = np.ascontiguousarray(dirty, np.complex64)
        if wgt is None:
            ier = L.dirty2msf_1(*args, d, epsilon, sigma, sign)
        else:
            ier = L.dirty2msf_2(*args, np.ascontiguousarray(wgt, np.float32), d, epsilon, sigma,
                                sign)
    check(ier, "dirty2vis")
    return ms


def vis2dirty_mixed(uvw, freq, ms, wgt, nxdirty, nydirty, fov, epsilon=1e-5, sigma=2.0,
                    sign=-1):
    """Mixed precision: FP64 uvw, complex64 vis/image, FP32 weights."""
    L = lib()
    dirty = np.zeros((nxdirty, nydirty), dtype=np.complex64)
    w = None if wgt is None else np.ascontiguousarray(wgt, np.float32)
    ier = L.ms2dirty_mixed(uvw.shape[0], nxdirty, nydirty, fov, float(np.ravel(freq)[0]),
                           np.ascontiguousarray(uvw, np.float64),
                           np.ascontiguousarray(ms, np.complex64).ravel(), w, dirty, epsilon,
                           sigma, sign)
    check(ier, "ms2dirty_mixed")
    return dirty


def dirty2vis_mixed(uvw, freq, dirty, wgt, fov, epsilon=1e-5, sigma=2.0, sign=-1):
    L = lib()
    nxdirty, nydirty = dirty.shape
    ms = np.zeros(uvw.shape[0], dtype=np.complex64)
    w = None if wgt is None else np.ascontiguousarray(wgt, np.float32)
    ier = L.dirty2ms_mixed(uvw.shape[0], nxdirty, nydirty, fov, float(np.ravel(freq)[0]),
                           np.ascontiguousarray(uvw, np.float64), ms, w,
                           np.ascontiguousarray(dirty, np.complex64), epsilon, sigma, sign)
    check(ier, "dirty2ms_mixed")
    return ms


def nufft2d1(x, y, c, N1, N2, epsilon=1e-6, sigma=2.0):
    """fk[N2, N1] = sum_j c_j exp(+i((k1) x_j + (k2) y_j)), x, y in [-pi, pi)."""
    L = lib()
    single = np.asarray(c).dtype == np.complex64
    if single:
        fk = np.zeros((N2, N1), dtype=np.complex64)
        ier = L.nufft2d1f(len(x), np.ascontiguousarray(x, np.float32),
                          np.ascontiguousarray(y, np.float32), np.ascontiguousarray(c, np.complex64),
                          N1, N2, fk, epsilon, sigma)
    else:
        fk = np.zeros((N2, N1), dtype=np.complex128)
        ier = L.nufft2d1(len(x), np.ascontiguousarray(x, np.float64),
                         np.ascontiguousarray(y, np.float64), np.ascontiguousarray(c, np.complex128),
                         N1, N2, fk, epsilon, sigma)
    check(ier, "nufft2d1")
    return fk


def nufft2d2(x, y, fk, epsilon=1e-6, sigma=2.0):
    """c_j = sum_k fk[k2, k1] exp(-i(k1 x_j + k2 y_j))."""
    L = lib()
    N2, N1 = fk.shape
    if fk.dtype == np.complex64:
        c = np.zeros(len(x), dtype=np.complex64)
        ier = L.nufft2d2f(len(x), np.ascontiguousarray(x, np.float32),
                          np.ascontiguousarray(y, np.float32), c, N1, N2,
                          np.ascontiguousarray(fk, np.complex64), epsilon, sigma)
    else:
        c = np.zeros(len(x), dtype=np.complex128)
        ier = L.nufft2d2(len(x), np.ascontiguousarray(x, np.float64),
                         np.ascontiguousarray(y, np.float64), c, N1, N2,
                         np.ascontiguousarray(fk, np.complex128), epsilon, sigma)
    check(ier, "nufft2d2")
    return c
