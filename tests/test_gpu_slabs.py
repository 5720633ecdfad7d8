"""The out-of-core W-stacking flow: a grid larger than the grid limit is
processed in slabs of w-planes through one buffer (the reference's mem_limit
pipeline, src/astro/ra_exec.cu:292-389 for ms2dirty and 520-613 for
dirty2ms, with partial_conv_3d_outputdriven_t1, src/fourier/conv.cu:1417-1542).

Each case runs the same call with the whole grid and with a grid limit that
forces several slabs (the last one possibly shorter), and checks the slabbed results against
the unslabbed ones and against the FP64 restatement of the reference
algorithm (oracle/).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
C_LIGHT = 299792458.0


def _rows(rng, nrow, f0, wmax_m):
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 300.0, (nrow, 2))
    uvw[:, 2] = rng.uniform(-wmax_m, wmax_m, nrow)
    return uvw


@pytest.mark.parametrize("prec,npix,fov,limit_mb", [
    (1, 200, 20.0, 3),     # complex128, non-power-of-two image
    (0, 256, 25.0, 4),     # complex64 data, FP64 uvw
    (0, 128, 40.0, 1),     # one or two planes per slab
])
def test_slabs_match_whole_grid_and_oracle(gpu, prec, npix, fov, limit_mb):
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(100 + npix)
    f0, nrow = 1.0e8, 20000
    eps = 1e-9 if prec == 1 else 1e-5
    cdt = np.complex128 if prec == 1 else np.complex64
    rdt = np.float64 if prec == 1 else np.float32
    uvw = _rows(rng, nrow, f0, 800.0)
    vis = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(cdt)
    wgt = rng.uniform(0.5, 1.5, nrow).astype(rdt)
    img = (rng.standard_normal((npix, npix)) + 1j * rng.standard_normal((npix, npix))).astype(cdt)
    tu, tv, tw, ti = (torch.from_numpy(a).cuda() for a in (uvw, vis, wgt, img))
    res = {}
    for limit in (0, limit_mb):
        ctx = Context(grid_limit_mb=limit)
        try:
            print("slabs prec=%d npix=%d limit=%d MiB" % (prec, npix, limit), flush=True)
            d = ctx.ms2dirty(tu, tv, tw, npix, npix, fov, f0, eps)
            s1 = ctx.last_slab()
            v = ctx.dirty2ms(tu, ti, tw, fov, f0, eps)
            s2 = ctx.last_slab()
            plan = ctx.last_plan()
            torch.cuda.synchronize()
            res[limit] = (d.cpu().numpy(), v.cpu().numpy(), s1, s2, plan)
        finally:
            ctx.close()
    d0, v0, s1, s2, plan = res[0]
    d1, v1, t1, t2, _ = res[limit_mb]
    nf3 = plan["nf3"]
    assert s1 == (0, 1) and s2 == (0, 1)
    # the limit forces several slabs, the last one possibly shorter
    assert t1 == t2 and t1[0] >= 1 and t1[1] >= 2 and t1[0] * t1[1] >= nf3 > t1[0] * (t1[1] - 1)
    plane_mb = plan["nf1"] * plan["nf2"] * (16 if prec == 1 else 8) / 2**20
    assert t1[0] * plane_mb <= limit_mb
    tol_same = 1e-13 if prec == 1 else 1e-6
    assert rel_l2(d1, d0) <= tol_same
    assert rel_l2(v1, v0) <= tol_same
    w64 = wgt.astype(np.float64)
    ref_d = O.ms2dirty(uvw, f0, vis.astype(np.complex128), w64, npix, npix, fov, eps,
                       single=prec == 0)
    ref_v = O.dirty2ms(uvw, f0, img.astype(np.complex128), w64, fov, eps, single=prec == 0)
    tol = eps if prec == 1 else 2e-6
    e1, e2 = rel_l2(d1, ref_d), rel_l2(v1, ref_v)
    print("slabs %s of %d planes: vs oracle %.2e %.2e" % (t1, nf3, e1, e2), flush=True)
    assert e1 <= tol and e2 <= tol


def test_slabs_adjoint_and_multichannel(gpu):
    """Adjointness of the slabbed pair and a multi-channel call in slabs
    against the same call with the whole grid."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(8)
    nrow, nchan, npix, fov, eps = 6000, 3, 160, 30.0, 1e-10
    freqs = np.array([1.0e8, 1.2e8, 1.5e8])
    uvw = _rows(rng, nrow, freqs[0], 600.0)
    vis = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    img = rng.standard_normal((npix, npix)) + 1j * rng.standard_normal((npix, npix))
    tu, tv, ti = (torch.from_numpy(a).cuda() for a in (uvw, vis, img))
    out = {}
    for limit in (0, 2):
        ctx = Context(grid_limit_mb=limit)
        try:
            d = ctx.ms2dirty(tu, tv, None, npix, npix, fov, freqs, eps)
            v = ctx.dirty2ms(tu, ti, None, fov, freqs, eps)
            torch.cuda.synchronize()
            out[limit] = (d.cpu().numpy(), v.cpu().numpy(), ctx.last_slab())
        finally:
            ctx.close()
    d, v, sl = out[2]
    assert sl[1] >= 2
    assert rel_l2(d, out[0][0]) <= 1e-13 and rel_l2(v, out[0][1]) <= 1e-13
    lhs = np.vdot(vis, v)
    rhs = np.vdot(d, img)
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)
