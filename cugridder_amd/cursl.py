"""NIFTY-compatible ms2dirty / dirty2ms over libcurafft.so.

Mirrors the reference binding python/curagridder/cursl.py:78-166: the same
function names, argument lists and conventions (sigma = 2, sign = -1, fov =
rad_pix_x * nxdirty in degrees, ms2dirty returns the real part of the complex
image, dirty2ms returns complex visibilities of shape (nrow, 1), dtype chooses
the double or single entry points, wgt may be None).
"""
import numpy as np

from ._lib import check, lib

__all__ = ["ms2dirty", "dirty2ms"]


def ms2dirty(uvw, freq, ms, wgt, nxdirty, nydirty, rad_pix_x, rad_pix_y, nx, ny, epsilon,
             do_wstacking, *args):
    """Image from visibilities by the improved W-stacking NUFFT.

    uvw (nrow, 3) metres; freq: frequencies (freq[0] is used); ms (nrow,) or
    (nrow, 1) complex; wgt (nrow,) or None; returns the real dirty image
    of shape (nxdirty, nydirty).  rad_pix_y, nx, ny and do_wstacking are
    accepted for NIFTY compatibility; the reference ignores them too
    (cursl.py:94-97).
    """
    L = lib()
    nrow = uvw.shape[0]
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    f0 = float(np.asarray(freq).ravel()[0])
    if ms.dtype == np.complex128:
        dirty = np.zeros((nxdirty, nydirty), dtype=np.complex128)
        u = np.ascontiguousarray(uvw, dtype=np.float64)
        m = np.ascontiguousarray(ms, dtype=np.complex128)
        if wgt is None:
            ier = L.ms2dirty_1(nrow, nxdirty, nydirty, fov, f0, u, m, dirty, epsilon, sigma, sign)
        else:
            w = np.ascontiguousarray(wgt, dtype=np.float64)
            ier = L.ms2dirty_2(nrow, nxdirty, nydirty, fov, f0, u, m, w, dirty, epsilon, sigma,
                               sign)
    else:
        dirty = np.zeros((nxdirty, nydirty), dtype=np.complex64)
        u = np.ascontiguousarray(uvw, dtype=np.float32)
        m = np.ascontiguousarray(ms, dtype=np.complex64)
        if wgt is None:
            ier = L.ms2dirtyf_1(nrow, nxdirty, nydirty, fov, f0, u, m, dirty, epsilon, sigma, sign)
        else:
            w = np.ascontiguousarray(wgt, dtype=np.float32)
            ier = L.ms2dirtyf_2(nrow, nxdirty, nydirty, fov, f0, u, m, w, dirty, epsilon, sigma,
                                sign)
    check(ier, "ms2dirty")
    dirty = np.reshape(dirty, [nxdirty, nydirty])
    return dirty.real


def dirty2ms(uvw, freq, dirty, wgt, rad_pix_x, rad_pix_y, nx, ny, epsilon, do_wstacking, *args):
    """Visibilities from a real image (adjoint of ms2dirty); returns (nrow, 1)."""
    L = lib()
    nrow = uvw.shape[0]
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    f0 = float(np.asarray(freq).ravel()[0])
    if dirty.dtype == np.float64:
        u = np.ascontiguousarray(uvw, dtype=np.float64)
        ms = np.zeros((nrow, 1), dtype=np.complex128)
        dirty1 = np.zeros(dirty.shape, dtype=np.complex128)
    else:
        u = np.ascontiguousarray(uvw, dtype=np.float32)
        ms = np.zeros((nrow, 1), dtype=np.complex64)
        dirty1 = np.zeros(dirty.shape, dtype=np.complex64)
    dirty1.real = dirty
    if dirty.dtype == np.float64:
        if wgt is None:
            ier = L.dirty2ms_1(nrow, nxdirty, nydirty, fov, f0, u, ms, dirty1, epsilon, sigma, sign)
        else:
            w = np.ascontiguousarray(wgt, dtype=np.float64)
            ier = L.dirty2ms_2(nrow, nxdirty, nydirty, fov, f0, u, ms, w, dirty1, epsilon, sigma,
                               sign)
    else:
        if wgt is None:
            ier = L.dirty2msf_1(nrow, nxdirty, nydirty, fov, f0, u, ms, dirty1, epsilon, sigma,
                                sign)
        else:
            w = np.ascontiguousarray(wgt, dtype=np.float32)
            ier = L.dirty2msf_2(nrow, nxdirty, nydirty, fov, f0, u, ms, w, dirty1, epsilon, sigma,
                                sign)
    check(ier, "dirty2ms")
    return ms
