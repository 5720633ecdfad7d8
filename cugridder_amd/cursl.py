"""NIFTY-compatible ms2dirty / dirty2ms over libcurafft.so.

Mirrors the reference binding python/curagridder/cursl.py:78-166: the same
function names, argument lists and conventions (sigma = 2, sign = -1, fov =
rad_pix_x * nxdirty in degrees, ms2dirty returns the real part of the complex
image, dirty2ms returns complex visibilities of shape (nrow, 1), dtype chooses
the double or single entry points, wgt may be None).

With several frequencies the reference binding still passes freq[0] only
(cursl.py:103); here ms and wgt of shape (nrow, nchan) are imaged with every
channel at its own frequency, as NIFTY does, through the multi-channel
entry points (complex64 data use FP64 uvw).  The reference also ignores
rad_pix_y (it derives the ny pixel size from rad_pix_x, cursl.py:99) and
do_wstacking (cursl.py:94-97); here both are honoured as NIFTY does, and the
reference's own entry points are used whenever the arguments mean the same
thing to both (do_wstacking true and rad_pix_y = rad_pix_x * nx / ny).
"""
import numpy as np

from ._lib import check, lib

__all__ = ["ms2dirty", "dirty2ms"]


def _geometry(nxdirty, nydirty, rad_pix_x, rad_pix_y, do_wstacking):
    """Checks the image geometry and says whether the reference's own entry
    points express it.  They take fov = rad_pix_x * nxdirty and derive the
    pixel size along ny as fov / nydirty (cugridder.cu:48-52), so a
    rad_pix_y that differs from rad_pix_x * nxdirty / nydirty (square
    pixels on a non-square image, say), or do_wstacking=False (plain 2-D
    gridding, w ignored and no 1/n, as NIFTY), go through ms2dirty_px /
    dirty2ms_px, which honour both as NIFTY does."""
    if int(nxdirty) != nxdirty or int(nydirty) != nydirty or nxdirty <= 0 or nydirty <= 0:
        raise ValueError("nxdirty and nydirty must be positive integers, got %r x %r"
                         % (nxdirty, nydirty))
    if not (np.isfinite(rad_pix_x) and rad_pix_x > 0):
        raise ValueError("rad_pix_x must be a positive pixel size in radians, got %r" % rad_pix_x)
    if rad_pix_y is None or rad_pix_y == 0:
        rad_pix_y = rad_pix_x * nxdirty / nydirty
    if not (np.isfinite(rad_pix_y) and rad_pix_y > 0):
        raise ValueError("rad_pix_y must be a positive pixel size in radians, got %r" % rad_pix_y)
    ref_y = rad_pix_x * nxdirty / nydirty
    reference = bool(do_wstacking) and abs(rad_pix_y - ref_y) <= 1e-12 * ref_y
    return float(rad_pix_y), reference


def _check_rows(uvw, a, name, nchan):
    if uvw.ndim != 2 or uvw.shape[1] != 3:
        raise ValueError("uvw must have shape (nrow, 3), got %s" % (uvw.shape,))
    if a is not None and np.size(a) != uvw.shape[0] * nchan:
        raise ValueError("%s must hold nrow * nchan = %d values, got shape %s"
                         % (name, uvw.shape[0] * nchan, np.shape(a)))


def _ms2dirty_px(uvw, fr, ms, wgt, nxdirty, nydirty, pix_x, pix_y, epsilon, do_wstacking):
    nrow, nchan = uvw.shape[0], fr.size
    double = ms.dtype == np.complex128
    cdt, rdt = (np.complex128, np.float64) if double else (np.complex64, np.float32)
    u = np.ascontiguousarray(uvw, dtype=np.float64)
    m = np.ascontiguousarray(np.reshape(ms, (nrow, nchan)), dtype=cdt)
    w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)), dtype=rdt)
    dirty = np.zeros((nxdirty, nydirty), dtype=cdt)
    ier = lib().ms2dirty_px(1 if double else 0, nrow, nchan, nxdirty, nydirty, pix_x, pix_y, fr,
                            u.ctypes.data, m.ctypes.data, None if w is None else w.ctypes.data,
                            dirty.ctypes.data, epsilon, 2.0, -1, 1 if do_wstacking else 0)
    check(ier, "ms2dirty")
    return dirty.real


def _dirty2ms_px(uvw, fr, dirty, wgt, pix_x, pix_y, epsilon, do_wstacking):
    nrow, nchan = uvw.shape[0], fr.size
    nxdirty, nydirty = dirty.shape
    double = dirty.dtype == np.float64
    cdt, rdt = (np.complex128, np.float64) if double else (np.complex64, np.float32)
    u = np.ascontiguousarray(uvw, dtype=np.float64)
    d = np.ascontiguousarray(dirty, dtype=cdt)
    w = None if wgt is None else np.ascontiguousarray(np.reshape(wgt, (nrow, nchan)), dtype=rdt)
    ms = np.zeros((nrow, nchan), dtype=cdt)
    ier = lib().dirty2ms_px(1 if double else 0, nrow, nchan, nxdirty, nydirty, pix_x, pix_y, fr,
                            u.ctypes.data, ms.ctypes.data, None if w is None else w.ctypes.data,
                            d.ctypes.data, epsilon, 2.0, -1, 1 if do_wstacking else 0)
    check(ier, "dirty2ms")
    return ms


def ms2dirty(uvw, freq, ms, wgt, nxdirty, nydirty, rad_pix_x, rad_pix_y, nx, ny, epsilon,
             do_wstacking, *args):
    """Image from visibilities by the improved W-stacking NUFFT.

    uvw (nrow, 3) metres; freq: nchan frequencies; ms (nrow,), (nrow, 1) or
    (nrow, nchan) complex; wgt of the same shape or None; returns the real
    dirty image of shape (nxdirty, nydirty).  rad_pix_y is honoured (the
    reference derives it from rad_pix_x, cursl.py:99); do_wstacking=False
    grids in 2-D, ignoring w, as NIFTY does (the reference ignores the flag);
    nx, ny and the trailing NIFTY arguments (nthreads, verbosity) are
    accepted and unused.
    """
    L = lib()
    uvw = np.asarray(uvw)
    nrow = uvw.shape[0] if uvw.ndim == 2 else -1
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    fr = np.ascontiguousarray(np.asarray(freq, dtype=np.float64).ravel())
    if fr.size == 0:
        raise ValueError("freq is empty")
    _check_rows(uvw, ms, "ms", fr.size)
    _check_rows(uvw, wgt, "wgt", fr.size)
    pix_y, reference = _geometry(nxdirty, nydirty, rad_pix_x, rad_pix_y, do_wstacking)
    if not reference:
        return _ms2dirty_px(uvw, fr, np.asarray(ms), wgt, int(nxdirty), int(nydirty),
                            float(rad_pix_x), pix_y, epsilon, do_wstacking)
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
    """Visibilities from a real image (adjoint of ms2dirty); returns (nrow, nchan).
    rad_pix_y and do_wstacking as in ms2dirty."""
    L = lib()
    uvw = np.asarray(uvw)
    dirty = np.asarray(dirty)
    if dirty.ndim != 2:
        raise ValueError("dirty must be a 2-D image, got shape %s" % (dirty.shape,))
    nxdirty, nydirty = dirty.shape[0], dirty.shape[1]
    sigma = 2
    fov = rad_pix_x * nxdirty * 180 / np.pi
    sign = -1
    fr = np.ascontiguousarray(np.asarray(freq, dtype=np.float64).ravel())
    if fr.size == 0:
        raise ValueError("freq is empty")
    _check_rows(uvw, None, "", fr.size)
    _check_rows(uvw, wgt, "wgt", fr.size)
    nrow = uvw.shape[0]
    pix_y, reference = _geometry(nxdirty, nydirty, rad_pix_x, rad_pix_y, do_wstacking)
    if not reference:
        return _dirty2ms_px(uvw, fr, dirty, wgt, float(rad_pix_x), pix_y, epsilon, do_wstacking)
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
