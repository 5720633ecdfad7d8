"""Device-resident execution: a reusable workspace bound to a HIP stream.

Inputs and outputs are torch tensors already on the GPU (torch is used only
for device memory and streams).  Every call goes through libcurafft.so's
cg_*_dev entry points.
"""
import ctypes

import numpy as np

from ._lib import check, lib

STAGES = ("plan", "sort", "records", "grid", "fft_rows", "fft_cols", "image")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check(t, name, dtypes, shape=None, numel=None):
    """Raises unless t is a contiguous GPU tensor of one of `dtypes` with
    the given shape (None entries match anything) or element count; the
    kernels read and write raw pointers, so a host tensor, a strided view or
    a short buffer must not reach them."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor, got %s" % (name, type(t).__name__))
    if not t.is_cuda:
        raise TypeError("%s must be on the GPU (got device %s)" % (name, t.device))
    if t.dtype not in dtypes:
        raise TypeError("%s must have dtype %s, got %s"
                        % (name, " or ".join(str(d) for d in dtypes), t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if shape is not None and (t.dim() != len(shape) or
                              any(s is not None and s != n for s, n in zip(shape, t.shape))):
        raise ValueError("%s must have shape %s, got %s"
                         % (name, tuple("*" if s is None else s for s in shape), tuple(t.shape)))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s must have %d elements, got %d" % (name, numel, t.numel()))
    return t


def _check_imaging(torch, uvw, data, wgt, nchan):
    """Argument checks shared by ms2dirty / dirty2ms: (prec, uvw_prec,
    wgt_prec) of the combinations the library implements: complex64 data
    with FP32 weights and FP64 or FP32 uvw, or complex128 data with FP64 uvw
    and weights."""
    prec = {torch.complex64: 0, torch.complex128: 1}.get(data.dtype)
    if prec is None:
        raise TypeError("vis / image must be complex64 or complex128, got %s" % data.dtype)
    uvw_types = (torch.float64, torch.float32) if prec == 0 else (torch.float64,)
    _check(uvw, "uvw", uvw_types, shape=(None, 3))
    nrow = uvw.shape[0]
    rdt = torch.float32 if prec == 0 else torch.float64
    if wgt is not None:
        _check(wgt, "wgt", (rdt,), numel=nrow * nchan)
    return prec, (1 if uvw.dtype == torch.float64 else 0), prec, nrow


class Context:
    """Wraps cg_ctx_create; `stream` is a torch.cuda.Stream (None = torch's
    current stream at construction)."""

    def __init__(self, stream=None, timing=False, **options):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("cugridder_amd.Context needs a GPU")
        self._torch = torch
        if stream is None:
            stream = torch.cuda.current_stream()
        self.stream = stream
        self._h = ctypes.c_void_p()
        check(lib().cg_ctx_create(ctypes.byref(self._h), ctypes.c_void_p(stream.cuda_stream)),
              "cg_ctx_create")
        if timing:
            self.set_option("timing", 1)
        for k, v in options.items():
            self.set_option(k, v)

    def set_option(self, name, value):
        check(lib().cg_ctx_set_option(self._h, name.encode(), int(value)), "set_option " + name)

    def set_kernel(self, kw, params, eps, deg=None):
        """Installs the learned kernel of support kw (1153 FP32 parameters),
        certified by the caller at relative error eps; eps <= 0 removes it.
        complex64 / mixed calls at epsilon >= eps whose ES kernel is wider
        use it."""
        from .learned import default_degree
        deg = default_degree(kw) if deg is None else deg
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        check(lib().cg_ctx_set_kernel(self._h, int(kw), int(deg), p, float(eps)), "set_kernel")

    def close(self):
        if self._h:
            lib().cg_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ imaging
    @staticmethod
    def _freqs(freq):
        f = np.ascontiguousarray(np.atleast_1d(np.asarray(freq, dtype=np.float64)).ravel())
        if f.size == 0:
            raise ValueError("freq is empty")
        return f

    @staticmethod
    def _pix(fov, nx, pix_y, do_wstacking):
        """(pix_x, pix_y) in radians when the call needs the NIFTY geometry
        (a pixel size along ny other than fov / ny, or plain 2-D gridding),
        else None."""
        if pix_y is None and do_wstacking:
            return None
        if not (fov > 0):
            raise ValueError("fov must be positive, got %r" % fov)
        pix_x = float(fov) * np.pi / 180.0 / nx
        if pix_y is not None and not (pix_y > 0):
            raise ValueError("pix_y must be a positive pixel size in radians, got %r" % pix_y)
        return pix_x, float(pix_y) if pix_y is not None else 0.0

    def ms2dirty(self, uvw, vis, wgt, nxdirty, nydirty, fov, freq, epsilon, sigma=2.0, sign=-1,
                 out=None, pix_y=None, do_wstacking=True):
        """Complex dirty image [nx, ny] from device visibilities.  freq is one
        frequency (vis [nrow]) or nchan of them (vis, wgt [nrow, nchan]); all
        channels are summed into the image.  fov (degrees) spans nxdirty;
        pix_y (radians) overrides the pixel size along ny (default fov /
        nydirty); do_wstacking=False grids in 2-D, ignoring w (no 1/n)."""
        torch = self._torch
        f = self._freqs(freq)
        if not isinstance(vis, torch.Tensor):
            raise TypeError("vis must be a torch tensor")
        prec, up, wp, nrow = _check_imaging(torch, uvw, vis, wgt, f.size)
        _check(vis, "vis", (vis.dtype,), numel=nrow * f.size)
        if out is None:
            out = torch.empty((nxdirty, nydirty), dtype=vis.dtype, device=vis.device)
        _check(out, "out", (vis.dtype,), shape=(nxdirty, nydirty))
        px = self._pix(fov, nxdirty, pix_y, do_wstacking)
        if px is None:
            ier = lib().cg_ms2dirty_mc_dev(self._h, prec, uvw.shape[0], f.size, nxdirty, nydirty,
                                           float(fov), f, _ptr(uvw), up, _ptr(vis), _ptr(wgt), wp,
                                           _ptr(out), float(epsilon), float(sigma), int(sign))
        else:
            py = px[1] if px[1] > 0 else float(fov) * np.pi / 180.0 / nydirty
            ier = lib().cg_ms2dirty_px_dev(self._h, prec, uvw.shape[0], f.size, nxdirty, nydirty,
                                           px[0], py, f, _ptr(uvw), up, _ptr(vis), _ptr(wgt), wp,
                                           _ptr(out), float(epsilon), float(sigma), int(sign),
                                           1 if do_wstacking else 0)
        check(ier, "ms2dirty")
        return out

    def dirty2ms(self, uvw, dirty, wgt, fov, freq, epsilon, sigma=2.0, sign=-1, out=None,
                 pix_y=None, do_wstacking=True):
        """Complex visibilities from a device image [nx, ny]: [nrow] for one
        frequency, [nrow, nchan] for nchan of them.  pix_y and do_wstacking
        as in ms2dirty."""
        torch = self._torch
        f = self._freqs(freq)
        if not isinstance(dirty, torch.Tensor):
            raise TypeError("dirty must be a torch tensor")
        prec, up, wp, nrow = _check_imaging(torch, uvw, dirty, wgt, f.size)
        _check(dirty, "dirty", (dirty.dtype,), shape=(None, None))
        nx, ny = dirty.shape
        if out is None:
            shape = (nrow,) if f.size == 1 else (nrow, f.size)
            out = torch.empty(shape, dtype=dirty.dtype, device=dirty.device)
        _check(out, "out", (dirty.dtype,), numel=nrow * f.size)
        px = self._pix(fov, nx, pix_y, do_wstacking)
        if px is None:
            ier = lib().cg_dirty2ms_mc_dev(self._h, prec, uvw.shape[0], f.size, nx, ny, float(fov),
                                           f, _ptr(uvw), up, _ptr(out), _ptr(wgt), wp, _ptr(dirty),
                                           float(epsilon), float(sigma), int(sign))
        else:
            py = px[1] if px[1] > 0 else float(fov) * np.pi / 180.0 / ny
            ier = lib().cg_dirty2ms_px_dev(self._h, prec, uvw.shape[0], f.size, nx, ny, px[0], py,
                                           f, _ptr(uvw), up, _ptr(out), _ptr(wgt), wp, _ptr(dirty),
                                           float(epsilon), float(sigma), int(sign),
                                           1 if do_wstacking else 0)
        check(ier, "dirty2ms")
        return out

    def batch(self, direction, items, epsilon, sigma=2.0, sign=-1, autotune=True, graph=True):
        """A Batch of independent channels on this context's stream."""
        return Batch(self, direction, items, epsilon, sigma, sign, autotune, graph)

    def _nufft_prec(self, x, y, data):
        torch = self._torch
        if not isinstance(data, torch.Tensor):
            raise TypeError("strengths / modes must be a torch tensor")
        prec = {torch.complex64: 0, torch.complex128: 1}.get(data.dtype)
        if prec is None:
            raise TypeError("strengths / modes must be complex64 or complex128, got %s" % data.dtype)
        rdt = torch.float32 if prec == 0 else torch.float64
        _check(x, "x", (rdt,), shape=(None,))
        _check(y, "y", (rdt,), numel=x.shape[0])
        return prec

    def nufft2d1(self, x, y, c, N1, N2, epsilon, sigma=2.0, out=None):
        """Type 1: modes [N2, N1] from strengths c at points (x, y) in
        [-pi, pi); x, y are float32 for complex64 c and float64 for
        complex128 c."""
        torch = self._torch
        prec = self._nufft_prec(x, y, c)
        _check(c, "c", (c.dtype,), numel=x.shape[0])
        if out is None:
            out = torch.empty((N2, N1), dtype=c.dtype, device=c.device)
        _check(out, "out", (c.dtype,), shape=(N2, N1))
        ier = lib().cg_nufft2d_dev(self._h, 1, prec, x.shape[0], _ptr(x), _ptr(y), _ptr(c), N1, N2,
                                   _ptr(out), float(epsilon), float(sigma))
        check(ier, "nufft2d1")
        return out

    def nufft2d2(self, x, y, fk, epsilon, sigma=2.0, out=None):
        """Type 2: strengths at points (x, y) from modes fk [N2, N1]."""
        torch = self._torch
        prec = self._nufft_prec(x, y, fk)
        _check(fk, "fk", (fk.dtype,), shape=(None, None))
        N2, N1 = fk.shape
        if out is None:
            out = torch.empty(x.shape[0], dtype=fk.dtype, device=fk.device)
        _check(out, "out", (fk.dtype,), numel=x.shape[0])
        ier = lib().cg_nufft2d_dev(self._h, 2, prec, x.shape[0], _ptr(x), _ptr(y), _ptr(out), N1,
                                   N2, _ptr(fk), float(epsilon), float(sigma))
        check(ier, "nufft2d2")
        return out

    # ------------------------------------------------------------ introspection
    def stage_times(self):
        """{'ms2dirty': {stage: ms}, 'dirty2ms': {...}, 'calls': (n1, n2)} summed;
        'grid_dense' is the MFMA kernel's part of each direction's 'grid'."""
        n = len(STAGES)
        out = np.zeros(2 * n + 2)
        calls = (ctypes.c_longlong * 2)()
        check(lib().cg_ctx_stage_times(self._h, out, calls), "stage_times")
        fw = dict(zip(STAGES, out[:n].tolist()))
        fw["grid_dense"] = float(out[2 * n])
        bw = dict(zip(STAGES, out[n:2 * n].tolist()))
        bw["grid_dense"] = float(out[2 * n + 1])
        return {"ms2dirty": fw, "dirty2ms": bw, "calls": (calls[0], calls[1])}

    def reset_times(self):
        lib().cg_ctx_reset_times(self._h)

    def last_plan(self):
        out = np.zeros(16)
        lib().cg_ctx_last_plan(self._h, out)
        keys = ("nf1", "nf2", "nf3", "kw", "beta", "es_c", "pix_x", "pix_y", "i_center",
                "i_half_width", "o_center", "o_half_width", "gamma", "h", "tile_x", "tile_y")
        d = dict(zip(keys, out.tolist()))
        for k in ("nf1", "nf2", "nf3", "kw", "tile_x", "tile_y"):
            d[k] = int(d[k])
        return d

    def last_band(self):
        """(first row, rows) of the grid rows the last W-stacking call's points
        reach along the column axis; rows = 0 when there is no band."""
        out = np.zeros(2, np.int32)
        lib().cg_ctx_last_band(self._h, out)
        return int(out[0]), int(out[1])

    def last_slab(self):
        """(w-planes per slab, slabs) of the last W-stacking call's grid
        buffer; (0, 1) when the whole grid was held at once."""
        out = np.zeros(2, np.int32)
        lib().cg_ctx_last_slab(self._h, out)
        return int(out[0]), int(out[1])

    # ------------------------------------------------------------ test hooks
    @staticmethod
    def spectrum_rows(n2):
        """Column-axis frequency held by each grid row after the forward
        column FFT (identity unless the axis is split, see cg_fft_col_split)."""
        na, nb = ctypes.c_int(), ctypes.c_int()
        split = lib().cg_fft_col_split(int(n2), ctypes.byref(na), ctypes.byref(nb))
        r = np.arange(n2)
        if split != 1:
            return r
        return r // nb.value + na.value * (r % nb.value)

    def grid_view(self, dtype):
        """The workspace w-plane grid of the last call as a torch tensor view
        [nf3, nf2, nf1] (copied)."""
        torch = self._torch
        p = ctypes.c_void_p()
        cells = ctypes.c_longlong()
        check(lib().cg_ctx_grid(self._h, ctypes.byref(p), ctypes.byref(cells)), "ctx_grid")
        plan = self.last_plan()
        n = cells.value
        itemsize = 8 if dtype == torch.complex64 else 16
        buf = torch.empty(n, dtype=dtype, device="cuda")
        _memcpy_d2d(buf.data_ptr(), p.value, n * itemsize)
        return buf.view(plan["nf3"], plan["nf2"], plan["nf1"])

    def tile_starts(self):
        """tile_start (int32, ntiles + 1) of the last call's sort."""
        torch = self._torch
        k, pm, ts = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n = (ctypes.c_longlong * 3)()
        check(lib().cg_ctx_sorted(self._h, ctypes.byref(k), ctypes.byref(pm), ctypes.byref(ts), n),
              "ctx_sorted")
        tst = torch.empty(n[1] + 1, dtype=torch.int32, device="cuda")
        _memcpy_d2d(tst.data_ptr(), ts.value, 4 * (n[1] + 1))
        self.key_shift = int(n[2])
        return tst

    def sorted_view(self):
        """(keys uint32-as-int64, perm int32, tile_start int32) of the last call."""
        torch = self._torch
        k, pm, ts = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        n = (ctypes.c_longlong * 3)()
        check(lib().cg_ctx_sorted(self._h, ctypes.byref(k), ctypes.byref(pm), ctypes.byref(ts), n),
              "ctx_sorted")
        m, nt = n[0], n[1]
        keys = torch.empty(m, dtype=torch.int32, device="cuda")
        perm = torch.empty(m, dtype=torch.int32, device="cuda")
        tst = torch.empty(nt + 1, dtype=torch.int32, device="cuda")
        if m:
            _memcpy_d2d(keys.data_ptr(), k.value, 4 * m)
            _memcpy_d2d(perm.data_ptr(), pm.value, 4 * m)
        _memcpy_d2d(tst.data_ptr(), ts.value, 4 * (nt + 1))
        self.key_shift = int(n[2])
        return keys.to(torch.int64) & 0xFFFFFFFF, perm, tst

    def fft2d(self, data, direction):
        """In-place unnormalised 2-D FFT of data [batch, n2, n1] (or [n2, n1])."""
        torch = self._torch
        prec = 0 if data.dtype == torch.complex64 else 1
        d = data if data.dim() == 3 else data.unsqueeze(0)
        check(lib().cg_fft2d_dev(self._h, prec, _ptr(d), d.shape[0], d.shape[1], d.shape[2],
                                 int(direction)), "fft2d")
        return data

    def sort_pairs(self, keys, vals, key_bits):
        """Stable in-place sort of int32 keys (as uint32) / int32 values."""
        check(lib().cg_sort_pairs_dev(self._h, _ptr(keys), _ptr(vals), keys.shape[0],
                                      int(key_bits)), "sort_pairs")
        return keys, vals


def _memcpy_d2d(dst, src, nbytes):
    import torch
    if nbytes == 0:
        return
    s = torch.cuda.current_stream()
    check(lib().cg_memcpy_dev(ctypes.c_void_p(dst), ctypes.c_void_p(src),
                              ctypes.c_void_p(s.cuda_stream), nbytes), "memcpy")
    s.synchronize()


class Batch:
    """Independent channels run as one launch (cg_batch_*): items are dicts
    with uvw (float64 [nrow, 3]), vis (complex64 [nrow]), wgt (float32 [nrow]
    or None), dirty (complex64 [nx, ny]), fov (deg) and freq (Hz) as device
    tensors / numbers.  direction "ms2dirty" writes every dirty, "dirty2ms"
    every vis.  The tensors must stay alive and unchanged in shape while the
    batch exists, and uvw must keep its contents: every item's plan (w-planes,
    the u / v ranges the FFT passes keep and the sort keys number) is fixed
    from the uvw present at creation.  vis, wgt and dirty contents may change
    between runs; rebuild the batch for new uvw."""

    def __init__(self, ctx, direction, items, epsilon, sigma=2.0, sign=-1, autotune=True,
                 graph=True):
        from ._lib import BATCH_AUTOTUNE, BATCH_GRAPH, BatchItem
        torch = ctx._torch
        self._ctx = ctx
        self.items = list(items)
        arr = (BatchItem * len(self.items))()
        for k, it in enumerate(self.items):
            _check(it["uvw"], "items[%d].uvw" % k, (torch.float64,), shape=(None, 3))
            m = it["uvw"].shape[0]
            _check(it["vis"], "items[%d].vis" % k, (torch.complex64,), numel=m)
            _check(it["dirty"], "items[%d].dirty" % k, (torch.complex64,), shape=(None, None))
            if it.get("wgt") is not None:
                _check(it["wgt"], "items[%d].wgt" % k, (torch.float32,), numel=m)
            nx, ny = it["dirty"].shape
            arr[k] = BatchItem(it["uvw"].shape[0], nx, ny, float(it["fov"]), float(it["freq"]),
                               _ptr(it["uvw"]), _ptr(it["vis"]), _ptr(it.get("wgt")),
                               _ptr(it["dirty"]))
        self._arr = arr
        flags = (BATCH_AUTOTUNE if autotune else 0) | (BATCH_GRAPH if graph else 0)
        self._h = ctypes.c_void_p()
        d = {"ms2dirty": 0, "dirty2ms": 1}[direction]
        check(lib().cg_batch_create(ctypes.byref(self._h), ctx._h, d, len(self.items), arr,
                                    float(epsilon), float(sigma), int(sign), flags),
              "cg_batch_create")

    def run(self):
        check(lib().cg_batch_run(self._h), "cg_batch_run")

    def info(self):
        """Per item: tile_grid, tile_degrid, dense threshold, tuned ms, nf3;
        and whether the batch runs as a graph."""
        out = np.zeros(5 * len(self.items))
        g = lib().cg_batch_info(self._h, out)
        return out.reshape(-1, 5), bool(g)

    def close(self):
        if self._h:
            lib().cg_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
