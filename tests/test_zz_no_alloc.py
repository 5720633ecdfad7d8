"""The allocation-free 2-D NUFFT entry point (cg_nufft2d_noalloc) against the
context entry point it mirrors (cg_nufft2d_dev), through the C-ABI.

Two calls in a row on the fixed workspace, the second with a smaller input
(the workspace keeps the first call's contents), for both types and both
precisions; a call whose workspace does not fit is refused with code 3
before anything runs.
"""
import ctypes

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _clustered(rng, M):
    x = rng.normal(0.7, 0.4, M)
    y = rng.normal(-1.1, 0.5, M)
    k = rng.random(M) < 0.3
    x[k] = rng.uniform(-np.pi, np.pi, k.sum())
    y[k] = rng.uniform(-np.pi, np.pi, k.sum())
    return (np.mod(x + np.pi, 2 * np.pi) - np.pi), (np.mod(y + np.pi, 2 * np.pi) - np.pi)


def _case(torch, rng, M, N1, N2, prec):
    rdt, cdt = (np.float32, np.complex64) if prec == 0 else (np.float64, np.complex128)
    x, y = _clustered(rng, M)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(cdt)
    fk = (rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))).astype(cdt)
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x.astype(rdt), y.astype(rdt), c, fk)]


def _run(torch, L, fn, h, typ, prec, x, y, c, fk, eps):
    """fn(h, ...) with fresh outputs; h is a context or a stream handle."""
    if typ == 1:
        out = torch.full_like(fk, 7.0)
        ier = fn(h, 1, prec, x.shape[0], ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                 ctypes.c_void_p(c.data_ptr()), fk.shape[1], fk.shape[0],
                 ctypes.c_void_p(out.data_ptr()), eps, 2.0)
    else:
        out = torch.full_like(c, -3.0)
        ier = fn(h, 2, prec, x.shape[0], ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                 ctypes.c_void_p(out.data_ptr()), fk.shape[1], fk.shape[0],
                 ctypes.c_void_p(fk.data_ptr()), eps, 2.0)
    torch.cuda.synchronize()
    return ier, out


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_context_entry(gpu, prec):
    torch = gpu
    from cugridder_amd import Context, lib
    L = lib()
    eps = 1e-5 if prec == 0 else 1e-9
    tol = 2e-6 if prec == 0 else 1e-12
    rng = np.random.default_rng(70 + prec)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ctx = Context()
    try:
        # a larger call, then a smaller one on the same (now dirty) workspace
        for (M, N1, N2) in ((200000, 512, 384), (3000, 96, 128)):
            x, y, c, fk = _case(torch, rng, M, N1, N2, prec)
            for typ in (1, 2):
                print("noalloc prec=%d type=%d M=%d %dx%d" % (prec, typ, M, N1, N2), flush=True)
                ier_a, a = _run(torch, L, L.cg_nufft2d_noalloc, stream, typ, prec, x, y, c, fk, eps)
                ier_b, b = _run(torch, L, L.cg_nufft2d_dev, ctx._h, typ, prec, x, y, c, fk, eps)
                assert ier_a == ier_b and ier_a in (0, 1), (ier_a, ier_b)
                e = rel_l2(a.cpu().numpy(), b.cpu().numpy())
                assert e <= tol, (typ, M, e)
    finally:
        ctx.close()


def test_noalloc_refuses_oversize(gpu):
    """A call whose workspace exceeds the fixed array returns 3 and leaves
    the output untouched (nothing was launched)."""
    torch = gpu
    from cugridder_amd import lib
    L = lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(5)
    x, y, c, fk = _case(torch, rng, 1000, 4096, 4096, 0)
    ier, out = _run(torch, L, L.cg_nufft2d_noalloc, stream, 1, 0, x, y, c, fk, 1e-5)
    assert ier == 3
    assert bool((out == 7.0).all())
    # and the fixed workspace still serves a call that fits
    x, y, c, fk = _case(torch, rng, 5000, 64, 64, 0)
    ier, a = _run(torch, L, L.cg_nufft2d_noalloc, stream, 2, 0, x, y, c, fk, 1e-5)
    assert ier in (0, 1)
    assert bool(torch.isfinite(torch.view_as_real(a)).all())
