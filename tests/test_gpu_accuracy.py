"""Accuracy of the headline and secondary workloads on the GPU, through the
device entry points, against the FP64 direct sums and the FP64 restatement
of the reference algorithm (oracle/).

* The full W-stacking geometry (8192^2, fov 3 deg, 1.4 GHz Earth-rotation
  tracks, |w| <= 5000 wavelengths, eps 1e-6: kw 7, nf 16384, 22 planes, the
  band-limited column pass, the dense MFMA gridder) on a row subsample and at
  the full 2^27 rows, against the direct sum on sampled pixels / rows, plus
  adjointness.  The method's own error at eps is a few eps against the
  direct sum (the reference prints it, python/curagridder/test_wst.py:91-95;
  test/w_s_gridder_test.cu:254-291); the complex64 path must not add more
  than a fraction of eps on top of what the FP64 restatement has on the
  same geometry at a size the oracle can run.
* vis2dirty / dirty2vis (sign +1), ported from the reference's
  python/curagridder/test_ws.py:34-66 (adjointness at rtol 1e-12), and
  compared with the oracle at sign +1.
* Batches: every item against the oracle (mixed sizes, a non-power-of-two
  image), and per-channel adjointness at the 64-channel workload's scale.
* The 2-D NUFFT at a power-of-two size (compile-time FFT plans) against the
  oracle.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT, as_float_abi, earth_rotation_rows, rel_l2

pytestmark = pytest.mark.gpu
C_LIGHT = 299792458.0


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ config 3

def _cfg3_run(torch, uvw, npix, fov, eps, vis, img, dense_min=64):
    from cugridder_amd import Context
    ctx = Context(timing=True, dense_min=dense_min)
    try:
        tu = torch.from_numpy(uvw).cuda() if isinstance(uvw, np.ndarray) else uvw
        tv = torch.from_numpy(vis).cuda() if isinstance(vis, np.ndarray) else vis
        ti = torch.from_numpy(img).cuda() if isinstance(img, np.ndarray) else img
        d = ctx.ms2dirty(tu, tv, None, npix, npix, fov, 1.4e9, eps)
        plan, band = ctx.last_plan(), ctx.last_band()
        v = ctx.dirty2ms(tu, ti, None, fov, 1.4e9, eps)
        torch.cuda.synchronize()
        st = ctx.stage_times()
    finally:
        ctx.close()
    return d, v, plan, band, st


def test_cfg3_geometry_subsample_vs_dft(gpu):
    """2^18 rows of the full W-stacking tracks on the full 8192^2 image at
    eps 1e-6 (complex64 data, FP64 uvw): the planned geometry is the full
    workload's (kw 7, nf 16384, 22 planes, band <= 2048 rows), dense bins go
    through the MFMA gridder, and both directions agree with the FP64 direct
    sum on 1000 sampled pixels / rows within 2 eps."""
    torch = gpu
    M, npix, fov, eps = 1 << 18, 8192, 3.0, 1e-6
    uvw = earth_rotation_rows(M, seed=3)
    rng = np.random.default_rng(33)
    vis = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    # a sparse image (the direct sum per row then costs its pixel count)
    npx = 1 << 15
    px = rng.choice(npix * npix, npx, replace=False).astype(np.int64)
    pv = (rng.standard_normal(npx) + 1j * rng.standard_normal(npx)).astype(np.complex64)
    img = np.zeros(npix * npix, np.complex64)
    img[px] = pv
    img = img.reshape(npix, npix)
    _say("cfg3 subsample: %d rows, %d^2" % (M, npix))
    d, v, plan, band, st = _cfg3_run(torch, uvw, npix, fov, eps, vis, img, dense_min=4)
    assert (plan["nf1"], plan["nf2"], plan["nf3"], plan["kw"]) == (16384, 16384, 22, 7)
    assert 0 < band[1] <= 2048
    assert st["ms2dirty"]["grid_dense"] > 0
    pix = fov * np.pi / 180 / npix
    sel = rng.choice(npix * npix, 1000, replace=False).astype(np.int64)
    t = O.dft_ms2dirty(uvw, 1.4e9, vis.astype(np.complex128), None, npix, npix, pix, pix,
                       pixels=sel)
    e1 = rel_l2(d.cpu().numpy().ravel()[sel], t)
    rows = rng.choice(M, 1000, replace=False).astype(np.int64)
    tv = O.dft_dirty2ms_sparse(uvw, 1.4e9, npix, npix, px, pv.astype(np.complex128), None, pix, pix,
                               rows=rows)
    e2 = rel_l2(v.cpu().numpy()[rows], tv)
    _say("cfg3 subsample vs direct sum: ms2dirty %.3e dirty2ms %.3e" % (e1, e2))
    assert e1 <= 2 * eps and e2 <= 2 * eps


def test_cfg3_pixel_size_vs_oracle(gpu):
    """The same tracks and pixel size on a 1024^2 image (fov 0.375 deg, eps
    1e-6, kw 7): the complex64 path against the FP64 restatement of the
    reference algorithm within eps on the whole image and every row, and the
    restatement's own error against the direct sum, which bounds what the
    full-size test above can reach."""
    torch = gpu
    M, npix, fov, eps = 1 << 17, 1024, 0.375, 1e-6
    uvw = earth_rotation_rows(M, seed=5)
    rng = np.random.default_rng(55)
    vis = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    img = (rng.standard_normal((npix, npix)) + 1j * rng.standard_normal((npix, npix))).astype(np.complex64)
    d, v, plan, band, st = _cfg3_run(torch, uvw, npix, fov, eps, vis, img, dense_min=4)
    assert plan["kw"] == 7 and band[1] > 0
    ref_d = O.ms2dirty(uvw, 1.4e9, vis.astype(np.complex128), None, npix, npix, fov, eps,
                       single=True)
    ref_v = O.dirty2ms(uvw, 1.4e9, img.astype(np.complex128), None, fov, eps, single=True)
    e1, e2 = rel_l2(d.cpu().numpy(), ref_d), rel_l2(v.cpu().numpy(), ref_v)
    pix = fov * np.pi / 180 / npix
    sel = rng.choice(npix * npix, 500, replace=False).astype(np.int64)
    t = O.dft_ms2dirty(uvw, 1.4e9, vis.astype(np.complex128), None, npix, npix, pix, pix, pixels=sel)
    e_or = rel_l2(ref_d.ravel()[sel], t)
    e_gpu = rel_l2(d.cpu().numpy().ravel()[sel], t)
    _say("cfg3 pixel size, 1024^2: vs restatement %.3e / %.3e; vs direct sum: restatement %.3e, "
         "GPU %.3e" % (e1, e2, e_or, e_gpu))
    assert e1 <= eps and e2 <= eps
    assert e_gpu <= e_or + 0.5 * eps


def test_cfg3_full_size_adjoint_and_dft(gpu):
    """The full W-stacking workload (2^27 rows, as the benchmark generates
    it): adjointness <vis, D(img)> = <G(vis), img> to 1e-5 relative, ms2dirty
    against the direct sum on 24 sampled pixels and dirty2ms on 500 sampled
    rows of a sparse image, within 2 eps."""
    torch = gpu
    sys.path.insert(0, ROOT)
    import bench
    W = bench.make_wstack_full(torch, "cuda:0")
    M, npix, eps = W["M"], W["nx"], W["eps"]
    from cugridder_amd import Context
    ctx = Context()
    try:
        d = ctx.ms2dirty(W["uvw"], W["vis"], None, npix, npix, W["fov"], W["freq"], eps)
        assert ctx.last_plan()["nf3"] == 22
        v = ctx.dirty2ms(W["uvw"], W["img"], None, W["fov"], W["freq"], eps)
        torch.cuda.synchronize()
        lhs = torch.vdot(W["vis"].to(torch.complex128), v.to(torch.complex128)).item()
        rhs = torch.vdot(d.to(torch.complex128).reshape(-1), W["img"].to(torch.complex128).reshape(-1)).item()
        _say("cfg3 full adjointness: %.3e" % (abs(lhs - rhs) / abs(rhs)))
        assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
        rng = np.random.default_rng(7)
        uvw = W["uvw"].cpu().numpy()
        vis = W["vis"].cpu().numpy().astype(np.complex128)
        pix = W["fov"] * np.pi / 180 / npix
        sel = rng.choice(npix * npix, 24, replace=False).astype(np.int64)
        t = O.dft_ms2dirty(uvw, W["freq"], vis, None, npix, npix, pix, pix, pixels=sel)
        e1 = rel_l2(d.cpu().numpy().ravel()[sel], t)
        del vis
        npx = 1 << 14
        px = rng.choice(npix * npix, npx, replace=False).astype(np.int64)
        pv = (rng.standard_normal(npx) + 1j * rng.standard_normal(npx)).astype(np.complex64)
        img = torch.zeros(npix * npix, dtype=torch.complex64, device="cuda")
        img[torch.from_numpy(px).cuda()] = torch.from_numpy(pv).cuda()
        v2 = ctx.dirty2ms(W["uvw"], img.view(npix, npix), None, W["fov"], W["freq"], eps)
        rows = rng.choice(M, 500, replace=False).astype(np.int64)
        tv = O.dft_dirty2ms_sparse(uvw, W["freq"], npix, npix, px, pv.astype(np.complex128), None,
                                   pix, pix, rows=rows)
        e2 = rel_l2(v2.cpu().numpy()[rows], tv)
        _say("cfg3 full vs direct sum: ms2dirty %.3e (24 pixels) dirty2ms %.3e (500 rows)" % (e1, e2))
        assert e1 <= 2 * eps and e2 <= 2 * eps
    finally:
        ctx.close()


# ------------------------------------------------- sign +1 (test_ws.py)

@pytest.mark.parametrize("nrow,npix,fov", [(1000, 512, 2.0), (10000, 512, 60.0)])
def test_vis2dirty_dirty2vis_sign_plus(gpu, nrow, npix, fov):
    """python/curagridder/test_ws.py:34-66 (vis2dirty / dirty2vis, sign +1,
    complex128, eps 1e-12): adjointness at rtol 1e-12, and both directions
    against the oracle called with sign +1 and against the direct sum with
    the +w phase of the reference's explicit_gridder (test_ws.py:13-30)."""
    from cugridder_amd import curagridder as cga
    eps, f0 = 1e-12, 1e9
    rng = np.random.default_rng(42)
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C_LIGHT)
    ms = (rng.random((nrow, 1)) - 0.5 + 1j * (rng.random((nrow, 1)) - 0.5))
    dirty = rng.random((npix, npix)) - 0.5 + 0j
    dirty2 = cga.vis2dirty(uvw, [f0], ms, None, np.zeros((npix, npix), np.complex128), fov, eps)
    ms2 = cga.dirty2vis(uvw, [f0], np.zeros((nrow, 1), np.complex128), None, dirty, fov, eps)
    np.testing.assert_allclose(np.vdot(ms, ms2).real, np.vdot(dirty2, dirty).real, rtol=1e-12)
    ref_d = O.ms2dirty(uvw, f0, ms.ravel(), None, npix, npix, fov, eps, sign=+1)
    ref_v = O.dirty2ms(uvw, f0, dirty, None, fov, eps, sign=+1)
    assert rel_l2(dirty2, ref_d) <= 1e-12
    assert rel_l2(ms2.ravel(), ref_v) <= 1e-12
    pix = fov * np.pi / 180 / npix
    sel = rng.choice(npix * npix, 300, replace=False).astype(np.int64)
    t = O.dft_ms2dirty(uvw, f0, ms.ravel(), None, npix, npix, pix, pix, sign=+1, pixels=sel)
    assert rel_l2(dirty2.ravel()[sel], t) <= 5e-11
    # the opposite w-sign is far off (the sign reaches the w-terms)
    t_wrong = O.dft_ms2dirty(uvw, f0, ms.ravel(), None, npix, npix, pix, pix, sign=-1, pixels=sel)
    assert rel_l2(dirty2.ravel()[sel], t_wrong) > 1e-6


@pytest.mark.parametrize("single", [False, True])
def test_sign_plus_float_and_mixed(gpu, single):
    """sign +1 through the float ABI (vis2dirty with complex64) and the
    mixed entry points, against the oracle at sign +1."""
    from cugridder_amd import curagridder as cga
    nrow, nx, ny, fov, f0, eps = 3000, 96, 128, 5.0, 1e9, 1e-5
    rng = np.random.default_rng(9)
    uvw = (rng.random((nrow, 3)) - 0.5) / (f0 / C_LIGHT)
    ms = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, nrow).astype(np.float32)
    img = (rng.standard_normal((nx, ny)) + 0j).astype(np.complex64)
    if single:
        d = cga.vis2dirty(uvw, [f0], ms, wgt, np.zeros((nx, ny), np.complex64), fov, eps)
        v = cga.dirty2vis(uvw, [f0], np.zeros(nrow, np.complex64), wgt, img, fov, eps)
        u_r, fov_r, f_r, e_r = as_float_abi(uvw), as_float_abi(fov), as_float_abi(f0), as_float_abi(eps)
    else:
        d = cga.vis2dirty_mixed(uvw, [f0], ms, wgt, nx, ny, fov, eps, sign=+1)
        v = cga.dirty2vis_mixed(uvw, [f0], img, wgt, fov, eps, sign=+1)
        u_r, fov_r, f_r, e_r = uvw, fov, f0, eps
    ref_d = O.ms2dirty(u_r, f_r, ms.astype(np.complex128), wgt.astype(np.float64), nx, ny, fov_r,
                       e_r, sign=+1, single=True)
    ref_v = O.dirty2ms(u_r, f_r, img.astype(np.complex128), wgt.astype(np.float64), fov_r, e_r,
                       sign=+1, single=True)
    assert rel_l2(d, ref_d) <= 2e-6
    assert rel_l2(v, ref_v) <= 2e-6


# ------------------------------------------------------------------ batches

def test_batch_items_vs_oracle(gpu):
    """Every item of a graph-captured, autotuned batch (mixed image sizes,
    including non-power-of-two ones, counts and frequencies, with and
    without weights) against the oracle in both directions."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(12)
    shapes = [(64, 64), (96, 128), (250, 64), (100, 100), (128, 128)]
    nrows = [4000, 1500, 9000, 2500, 6000]
    eps = 1e-5
    items, host = [], []
    for k, ((nx, ny), m) in enumerate(zip(shapes, nrows)):
        uvw = np.empty((m, 3))
        uvw[:, :2] = rng.normal(0, 250, (m, 2))
        uvw[:, 2] = rng.uniform(-60, 60, m)
        vis = (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64)
        wgt = rng.uniform(0.5, 1.5, m).astype(np.float32) if k % 2 else None
        img = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
        fov, f = float(rng.uniform(1.0, 4.0)), float(rng.uniform(1e8, 3e8))
        host.append((uvw, vis, wgt, img, fov, f))
        items.append(dict(uvw=torch.from_numpy(uvw).cuda(), vis=torch.from_numpy(vis).cuda(),
                          wgt=None if wgt is None else torch.from_numpy(wgt).cuda(),
                          img=torch.from_numpy(img).cuda(),
                          dirty=torch.zeros((nx, ny), dtype=torch.complex64, device="cuda"),
                          out=torch.zeros(m, dtype=torch.complex64, device="cuda"), fov=fov, freq=f))
    ctx = Context()
    try:
        b1 = ctx.batch("ms2dirty", items, eps)
        b2 = ctx.batch("dirty2ms", [dict(it, dirty=it["img"], vis=it["out"]) for it in items], eps)
        for rep in range(2):
            for it in items:
                it["dirty"].fill_(3.0)
                it["out"].fill_(-1.0)
            b1.run()
            b2.run()
            torch.cuda.synchronize()
            for k, (uvw, vis, wgt, img, fov, f) in enumerate(host):
                nx, ny = img.shape
                w64 = None if wgt is None else wgt.astype(np.float64)
                ref_d = O.ms2dirty(uvw, f, vis.astype(np.complex128), w64, nx, ny, fov, eps,
                                   single=True)
                ref_v = O.dirty2ms(uvw, f, img.astype(np.complex128), w64, fov, eps, single=True)
                e1 = rel_l2(items[k]["dirty"].cpu().numpy(), ref_d)
                e2 = rel_l2(items[k]["out"].cpu().numpy(), ref_v)
                assert e1 <= 2e-6 and e2 <= 2e-6, (rep, k, e1, e2)
        b1.close()
        b2.close()
    finally:
        ctx.close()


def test_batch_cfg4_scale_adjointness(gpu):
    """The 64-channel workload as the benchmark builds it (2^18 .. 2^23 rows,
    images 512^2 .. 4096^2, 1.0 .. 1.6 GHz): every channel's batched
    ms2dirty and dirty2ms satisfy adjointness to 1e-5 relative."""
    torch = gpu
    sys.path.insert(0, ROOT)
    import bench
    from cugridder_amd import Context
    W = bench.make_batch(torch, "cuda:0")
    ctx = Context()
    try:
        fw = [dict(it) for it in W["items"]]
        bw = [dict(it, dirty=it["img"], vis=it["out"]) for it in W["items"]]
        b1 = ctx.batch("ms2dirty", fw, W["eps"])
        b2 = ctx.batch("dirty2ms", bw, W["eps"])
        b1.run()
        b2.run()
        torch.cuda.synchronize()
        worst = 0.0
        for it in W["items"]:
            c128 = torch.complex128
            lhs = torch.vdot(it["vis"].to(c128), it["out"].to(c128)).item()
            rhs = torch.vdot(it["dirty"].to(c128).reshape(-1), it["img"].to(c128).reshape(-1)).item()
            worst = max(worst, abs(lhs - rhs) / abs(rhs))
        _say("cfg4 batch: worst channel adjointness %.3e over %d channels" % (worst, len(W["items"])))
        assert worst <= 1e-5
        b1.close()
        b2.close()
    finally:
        ctx.close()


# --------------------------------------------------------- 2-D NUFFT, 2^k

@pytest.mark.parametrize("N1,N2", [(1024, 1024), (2048, 1024)])
def test_nufft2d_power_of_two_vs_oracle(gpu, N1, N2):
    """complex64 2-D NUFFT with clustered points on power-of-two grids (the
    compile-time row and column FFT plans) against the FP32-ABI oracle."""
    torch = gpu
    from cugridder_amd import Context
    from tests.test_gpu_parity import _clustered_xy
    rng = np.random.default_rng(N1 + N2)
    M, eps = 200000, 1e-5
    x, y = _clustered_xy(rng, M)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    fk = (rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))).astype(np.complex64)
    ctx = Context()
    try:
        tx, ty, tc, tf = (torch.from_numpy(a).cuda() for a in (x, y, c, fk))
        f1 = ctx.nufft2d1(tx, ty, tc, N1, N2, eps).cpu().numpy()
        assert ctx.last_plan()["nf1"] == 2 * N1
        c2 = ctx.nufft2d2(tx, ty, tf, eps).cpu().numpy()
    finally:
        ctx.close()
    ref1 = O.nufft2d1(x, y, c, N1, N2, as_float_abi(eps), single=True)
    ref2 = O.nufft2d2(x, y, fk, as_float_abi(eps), single=True)
    e1, e2 = rel_l2(f1, ref1), rel_l2(c2, ref2)
    _say("nufft2d %dx%d vs oracle: type 1 %.3e type 2 %.3e" % (N1, N2, e1, e2))
    assert e1 <= eps and e2 <= eps
