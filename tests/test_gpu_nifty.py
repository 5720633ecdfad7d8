"""The NIFTY surface of cursl on the GPU: do_wstacking=False (plain 2-D
gridding, w ignored, no 1/n) and a pixel size along ny other than the
reference's fov / ny (python/curagridder/cursl.py:78-166 accepts both
arguments and ignores them; NIFTY honours them).

Plain 2-D is checked against the oracle's 2-D NUFFT restatement on the
scaled (v, u) coordinates (conv_2d_nputsdriven + deconv_2d, the same
kernel, grid and corrections) and against the direct 2-D sum; separate pixel
sizes against the W-stacking restatement planned with that pixel size and
against the direct sum with the w-term.  Adjointness as the reference
asserts it (python/curagridder/test_wst.py:84-89)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, wst_case

pytestmark = pytest.mark.gpu


def _floor(single):
    return 2e-6 if single else 1e-13


@pytest.mark.parametrize("nrow,nx,ny,fov", [(300, 64, 96, 5.0), (2000, 128, 128, 2.0),
                                            (500, 250, 64, 20.0)])
@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("use_wgt", [False, True])
def test_plain_2d(gpu, nrow, nx, ny, fov, single, use_wgt):
    import cugridder_amd as cg
    eps = 1e-5 if single else 1e-10
    k = wst_case(nrow, nx, ny, fov, use_wgt=use_wgt)
    ms, dirty, wgt = k["ms"], k["dirty"], k["wgt"]
    if single:
        ms, dirty = ms.astype("c8"), dirty.astype("f4")
        wgt = None if wgt is None else wgt.astype("f4")
    xp, yp = k["xpix"], k["ypix"]
    d = cg.ms2dirty(k["uvw"], [k["f0"]], ms, wgt, nx, ny, xp, yp, 0, 0, eps, False).astype("f8")
    v = cg.dirty2ms(k["uvw"], [k["f0"]], dirty, wgt, xp, yp, 0, 0, eps, False).ravel()
    tol = 5e-5 if single else 1e-12
    np.testing.assert_allclose(np.vdot(ms, v).real, np.vdot(d, dirty), rtol=tol)
    w64 = None if wgt is None else wgt.astype("f8")
    ref_d = O.ms2dirty_2d(k["uvw"], k["f0"], ms, w64, nx, ny, xp, yp, eps, single=single)
    ref_v = O.dirty2ms_2d(k["uvw"], k["f0"], dirty.astype("c16"), w64, xp, yp, eps, single=single)
    assert rel_l2(d, ref_d.real) <= max(eps, _floor(single))
    assert rel_l2(v, ref_v) <= max(eps, _floor(single))
    # the direct 2-D sum (no w-term, no 1/n)
    x, y = O.uv_coords_2d(k["uvw"], k["f0"], xp, yp)
    c = ms.astype("c16") * (1.0 if w64 is None else w64)
    truth = O.dft2d1(x, y, c, ny, nx).reshape(nx, ny)
    assert rel_l2(d, truth.real) <= 5 * max(eps, 1e-6 if single else 0)
    # the w-stacking image of the same data differs (the w-term and 1/n)
    if fov >= 5:
        d_w = cg.ms2dirty(k["uvw"], [k["f0"]], ms, wgt, nx, ny, xp, yp, 0, 0, eps, True)
        assert rel_l2(d_w, d) > 1e-4


@pytest.mark.parametrize("nx,ny", [(64, 128), (250, 64)])
@pytest.mark.parametrize("single", [False, True])
def test_square_pixels_non_square_image(gpu, nx, ny, single):
    """rad_pix_y = rad_pix_x on an nx != ny image: the reference would use
    rad_pix_x * nx / ny along ny; here the given pixel size is honoured."""
    import cugridder_amd as cg
    nrow, fov, eps = 1500, 8.0, (1e-5 if single else 1e-10)
    k = wst_case(nrow, nx, ny, fov, use_wgt=True)
    xp = k["xpix"]
    yp = xp
    ms, dirty, wgt = k["ms"], k["dirty"], k["wgt"]
    if single:
        ms, dirty, wgt = ms.astype("c8"), dirty.astype("f4"), wgt.astype("f4")
    d = cg.ms2dirty(k["uvw"], [k["f0"]], ms, wgt, nx, ny, xp, yp, 0, 0, eps, True).astype("f8")
    v = cg.dirty2ms(k["uvw"], [k["f0"]], dirty, wgt, xp, yp, 0, 0, eps, True).ravel()
    tol = 5e-5 if single else 1e-12
    np.testing.assert_allclose(np.vdot(ms, v).real, np.vdot(d, dirty), rtol=tol)
    fov_deg = xp * nx * 180 / np.pi
    w64 = wgt.astype("f8")
    ref_d = O.ms2dirty(k["uvw"], k["f0"], ms.astype("c16"), w64, nx, ny, fov_deg, eps,
                       single=single, pix_y=yp).real
    ref_v = O.dirty2ms(k["uvw"], k["f0"], dirty.astype("c16"), w64, fov_deg, eps,
                       single=single, pix_y=yp)
    assert rel_l2(d, ref_d) <= max(eps, _floor(single))
    assert rel_l2(v, ref_v) <= max(eps, _floor(single))
    truth = O.dft_ms2dirty(k["uvw"], k["f0"], ms.astype("c16"), w64, nx, ny, xp, yp)
    assert rel_l2(d.ravel(), truth.real) <= 5 * max(eps, 1e-6 if single else 0)
    # the reference's derivation of the ny pixel size gives another image
    d_ref_geom = cg.ms2dirty(k["uvw"], [k["f0"]], ms, wgt, nx, ny, xp, xp * nx / ny, 0, 0, eps,
                             True)
    assert rel_l2(d_ref_geom, d) > 1e-3


@pytest.mark.parametrize("single", [False, True])
def test_context_plain_2d_and_pix_y(gpu, single):
    """The device entry points with the same geometry options, against the
    host ones."""
    import cugridder_amd as cg
    from cugridder_amd import Context
    torch = gpu
    nrow, nx, ny, fov = 3000, 96, 160, 4.0
    eps = 1e-5 if single else 1e-10
    k = wst_case(nrow, nx, ny, fov, use_wgt=True)
    cdt, rdt = ("c8", "f4") if single else ("c16", "f8")
    ms, wgt = k["ms"].astype(cdt), k["wgt"].astype(rdt)
    img = (k["dirty"] + 0.25j).astype(cdt)
    xp = k["xpix"]
    ctx = Context()
    try:
        tu = torch.from_numpy(k["uvw"]).cuda()
        tv, tw, ti = (torch.from_numpy(a).cuda() for a in (ms, wgt, img))
        for pix_y, wst in ((None, False), (xp, True), (0.7 * xp, False)):
            d = ctx.ms2dirty(tu, tv, tw, nx, ny, fov, k["f0"], eps, pix_y=pix_y,
                             do_wstacking=wst).cpu().numpy()
            v = ctx.dirty2ms(tu, ti, tw, fov, k["f0"], eps, pix_y=pix_y,
                             do_wstacking=wst).cpu().numpy()
            yp = pix_y if pix_y is not None else fov * np.pi / 180 / ny
            if wst:
                ref_d = O.ms2dirty(k["uvw"], k["f0"], ms, wgt, nx, ny, fov, eps, single=single,
                                   pix_y=yp)
                ref_v = O.dirty2ms(k["uvw"], k["f0"], img, wgt, fov, eps, single=single, pix_y=yp)
            else:
                ref_d = O.ms2dirty_2d(k["uvw"], k["f0"], ms, wgt, nx, ny, xp, yp, eps,
                                      single=single)
                ref_v = O.dirty2ms_2d(k["uvw"], k["f0"], img, wgt, xp, yp, eps, single=single)
            assert rel_l2(d, ref_d) <= max(eps, _floor(single)), (pix_y, wst)
            assert rel_l2(v, ref_v) <= max(eps, _floor(single)), (pix_y, wst)
    finally:
        ctx.close()
