"""NIFTY-compatible ms2dirty / dirty2ms over libcurafft.so.

Mirrors the reference binding python/curagridder/cursl.py:78-166: the same
function names, argument lists and conventions (sigma = 2, sign = -1, fov =
rad_pix_x * nxdirty in degrees, ms2dirty returns the real part of the complex
image, dirty2ms returns complex visibilities of shape (nrow, 1), dtype chooses
the double or single entry points, wgt may be None).

With several frequencies the reference binding still passes freq[0] only
(cursl.py:103); here ms and wgt of shape (nrow, nchan) are imaged with every
channel at its own frequency, as NIFTY does, through the multi-channel
entry points (complex64 data use FP64 uvw).
"""
import numpy as np

from ._lib import check, lib

__all__ = ["ms2dirty", "dirty2ms"]


def ms2dirty(uvw, freq, ms, wgt, nxdirty, nydirty, rad_pix_x, rad_pix_y, nx, ny, epsilon,
             do_wstacking, *args):
    """Image from visibilities by the improved W-stacking NUFFT.

    uvw (nrow, 3) metres; freq: nchan frequencies; ms (nrow,), (nrow, 1) or
    (nrow, nchan) complex; wgt of the same shape or None; returns the real
    dirty image of shape (nxdirty, nydirty).  rad_pix_y, nx, ny and
    do_wstacking are accepted for NIFTY compatibility; the reference ignores
    them too (cursl.py:94-97).
    """
    L = lib()
    nrow = uvw.shape[0]
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    fr = np.ascontiguousarray(np.asarray(freq, dtype=np.float64).ravel())
    f0 = float(fr[0])
    if fr.size > 1:
        nchan = fr.size
        u = np.ascontiguousarray(uvw, dtype=np.float64)
        if ms.dtype == np.complex128:
            dirty = np.zeros((nxdirty, nydirty), dtype=np.complex128)
            m = np.ascontiguousarray(np.reshape(ms, (nrow, nchan)), dtype=np.complex128)
            w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)),
                                                              dtype=np.float64)
            ier = L.ms2dirty_mc_2(nrow, nchan, nxdirty, nydirty, fov, fr, u, m, w, dirty, epsilon,
                                  sigma, sign)
        else:
            dirty = np.zeros((nxdirty, nydirty), dtype=np.complex64)
            m = np.ascontiguousarray(np.reshape(ms, (nrow, nchan)), dtype=np.complex64)
            w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)),
                                                              dtype=np.float32)
            ier = L.ms2dirty_mc_mixed(nrow, nchan, nxdirty, nydirty, fov, fr, u, m, w, dirty,
                                      epsilon, sigma, sign)
        check(ier, "ms2dirty")
        return dirty.real
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
    """Visibilities from a real image (adjoint of ms2dirty); returns (nrow, nchan)."""
    L = lib()
    nrow = uvw.shape[0]
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    fr = np.ascontiguousarray(np.asarray(freq, dtype=np.float64).ravel())
    f0 = float(fr[0])
    if fr.size > 1:
        nchan = fr.size
        u = np.ascontiguousarray(uvw, dtype=np.float64)
        if dirty.dtype == np.float64:
            ms = np.zeros((nrow, nchan), dtype=np.complex128)
            d = np.ascontiguousarray(dirty, dtype=np.complex128)
            w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)),
                                                              dtype=np.float64)
            ier = L.dirty2ms_mc_2(nrow, nchan, nxdirty, nydirty, fov, fr, u, ms, w, d, epsilon,
                                  sigma, sign)
        else:
            ms = np.zeros((nrow, nchan), dtype=np.complex64)
            d = np.ascontiguousarray(dirty, dtype=np.complex64)
            w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)),
                                                              dtype=np.float32)
            ier = L.dirty2ms_mc_mixed(nrow, nchan, nxdirty, nydirty, fov, fr, u, ms, w, d,
                                      epsilon, sigma, sign)
        check(ier, "dirty2ms")
        return ms
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
