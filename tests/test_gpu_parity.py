"""End-to-end parity on the GPU through the C-ABI (host entry points used by
the reference's ctypes binding, and the device-pointer entry points).

Criteria: against the FP64 restatement of the reference algorithm (same
inputs, same kw/beta/nf/nf3), relative L2 <= eps (FP64) or <= max(eps, 2e-6)
(FP32 arithmetic); against the exact direct sum, the method's own accuracy
(5 eps for ms2dirty, 100 eps for dirty2ms on few rows; see test_oracle.py);
adjointness as the reference asserts it (rtol 1e-12 FP64, 5e-5 FP32,
python/curagridder/test_wst.py:84-89)."""
import glob
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import as_float_abi, rel_l2, wst_case

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _floor(single):
    return 2e-6 if single else 1e-13


def _assert_adjoint(lhs, rhs, ms, m2, tol):
    """<ms, D(dirty)> = <G(ms), dirty> to rtol tol, as the reference asserts
    it (python/curagridder/test_wst.py:84-89).  When the two sums cancel to
    far below the Cauchy-Schwarz scale |ms| |m2| (some of the crossed image
    shapes draw such inputs), the rounding of the single-precision results
    is measured against one hundredth of that scale instead of the
    cancelled value."""
    scale = max(abs(rhs), 1e-2 * np.linalg.norm(ms) * np.linalg.norm(m2))
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs, scale)


# reference parameter grid, python/curagridder/test_wst.py:36-43: nrow, nxdirty
# crossed with nydirty, fov, eps, precision, weights
SWEEP = list(itertools.product((2, 27, 100), tuple(itertools.product((32, 64, 16), (128, 250, 64))),
                               (1, 10, 20), (2e-1, 5e-3, 5e-5, 5e-7, 5e-12), (False, True),
                               (False, True)))


@pytest.mark.parametrize("nrow,shape,fov,eps,single,use_wgt", SWEEP)
def test_wst_sweep(gpu, nrow, shape, fov, eps, single, use_wgt):
    if single and eps < 5e-6:
        pytest.skip("single precision below 5e-6 (as the reference's test_wst.py:52-53)")
    import cugridder_amd as cg
    nx, ny = shape
    k = wst_case(nrow, nx, ny, fov, use_wgt=use_wgt)
    ms, dirty, wgt = k["ms"], k["dirty"], k["wgt"]
    if single:
        ms = ms.astype("c8")
        dirty = dirty.astype("f4")
        wgt = wgt.astype("f4") if wgt is not None else None
    d2 = cg.ms2dirty(k["uvw"], [k["f0"]], ms, wgt, nx, ny, k["xpix"], k["ypix"], 0, 0, eps,
                     True).astype("f8")
    m2 = cg.dirty2ms(k["uvw"], [k["f0"]], dirty, wgt, k["xpix"], k["ypix"], 0, 0, eps,
                     True).astype("c16").ravel()
    tol = 5e-5 if single else 1e-12
    _assert_adjoint(np.vdot(ms, m2).real, np.vdot(d2, dirty), ms, m2, tol)
    # the oracle gets exactly the inputs the entry point received
    w64 = None if wgt is None else wgt.astype("f8")
    if single:
        uvw_r, fov_r, f0_r, eps_r = (as_float_abi(k["uvw"]), as_float_abi(k["xpix"] * nx * 180 / np.pi),
                                     as_float_abi(k["f0"]), as_float_abi(eps))
    else:
        uvw_r, fov_r, f0_r, eps_r = k["uvw"], k["xpix"] * nx * 180 / np.pi, k["f0"], eps
    ref_d = O.ms2dirty(uvw_r, f0_r, ms.astype("c16"), w64, nx, ny, fov_r, eps_r,
                       single=single).real
    ref_v = O.dirty2ms(uvw_r, f0_r, dirty.astype("c16"), w64, fov_r, eps_r, single=single)
    assert rel_l2(d2, ref_d) <= max(eps, _floor(single))
    assert rel_l2(m2, ref_v) <= max(eps, _floor(single))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "wstack_*.npz"))))
@pytest.mark.parametrize("prec", ["double", "single", "mixed"])
def test_golden_wstack(gpu, path, prec):
    import cugridder_amd as cg
    from cugridder_amd import curagridder as cga
    g = np.load(path)
    nx, ny = g["dirty"].shape
    f0, fov = float(g["freq"]), float(g["fov"])
    eps = 1e-9 if prec == "double" else 1e-5
    L = cg.lib()
    if prec == "double":
        d = np.zeros((nx, ny), np.complex128)
        assert L.ms2dirty_2(len(g["vis"]), nx, ny, fov, f0, g["uvw"], g["vis"], g["wgt"], d, eps,
                            2.0, -1) == 0
        v = np.zeros(len(g["vis"]), np.complex128)
        assert L.dirty2ms_2(len(g["vis"]), nx, ny, fov, f0, g["uvw"], v, g["wgt"], g["dirty"], eps,
                            2.0, -1) == 0
    elif prec == "single":
        d = np.zeros((nx, ny), np.complex64)
        assert L.ms2dirtyf_2(len(g["vis"]), nx, ny, fov, f0, g["uvw"].astype("f4"),
                             g["vis"].astype("c8"), g["wgt"].astype("f4"), d, eps, 2.0, -1) == 0
        v = np.zeros(len(g["vis"]), np.complex64)
        assert L.dirty2msf_2(len(g["vis"]), nx, ny, fov, f0, g["uvw"].astype("f4"), v,
                             g["wgt"].astype("f4"), g["dirty"].astype("c8"), eps, 2.0, -1) == 0
    else:
        d = cga.vis2dirty_mixed(g["uvw"], [f0], g["vis"], g["wgt"], nx, ny, fov, eps)
        v = cga.dirty2vis_mixed(g["uvw"], [f0], g["dirty"], g["wgt"], fov, eps)
    single = prec != "double"
    assert rel_l2(d, g["dft_dirty"]) < 5 * eps + (3e-6 if single else 0)
    assert rel_l2(v, g["dft_vis"]) < 100 * eps + (3e-6 if single else 0)


@pytest.mark.parametrize("prec", ["double", "single"])
def test_golden_nufft2d(gpu, prec):
    from cugridder_amd import curagridder as cga
    g = np.load(os.path.join(GOLDEN, "nufft2d_500_48x32.npz"))
    N2, N1 = g["fk"].shape
    eps = 1e-9 if prec == "double" else 1e-5
    cdt = np.complex128 if prec == "double" else np.complex64
    fdt = np.float64 if prec == "double" else np.float32
    f = cga.nufft2d1(g["x"].astype(fdt), g["y"].astype(fdt), g["c"].astype(cdt), N1, N2, eps)
    c = cga.nufft2d2(g["x"].astype(fdt), g["y"].astype(fdt), g["fk"].astype(cdt), eps)
    floor = 0 if prec == "double" else 3e-6
    assert rel_l2(f, g["dft_type1"]) < 10 * eps + floor
    assert rel_l2(c, g["dft_type2"]) < 10 * eps + floor
    if prec == "double":
        ref1 = O.nufft2d1(g["x"], g["y"], g["c"], N1, N2, eps)
    else:
        ref1 = O.nufft2d1(as_float_abi(g["x"]), as_float_abi(g["y"]), as_float_abi(g["c"]), N1, N2,
                          as_float_abi(eps), single=True)
    assert rel_l2(f, ref1) <= max(eps, _floor(prec != "double"))


def _clustered_xy(rng, m):
    """80% of the points from 4 wrapped Gaussians (std 0.3), 20% uniform,
    in [-pi, pi)^2 (the shape of the 2-D NUFFT benchmark workload)."""
    nb = int(0.8 * m)
    means = rng.uniform(-np.pi, np.pi, (4, 2))
    p = np.concatenate([means[rng.integers(0, 4, nb)] + 0.3 * rng.standard_normal((nb, 2)),
                        rng.uniform(-np.pi, np.pi, (m - nb, 2))])
    p = np.remainder(p + np.pi, 2 * np.pi) - np.pi
    return p[:, 0].astype(np.float32), p[:, 1].astype(np.float32)


@pytest.mark.parametrize("tile,N1,N2,eps", [(8, 96, 64, 1e-5), (16, 96, 64, 1e-5), (32, 96, 64, 1e-5),
                                             (64, 96, 64, 1e-5), (32, 32, 48, 2e-6), (16, 40, 24, 1e-3)])
def test_nufft2d_tiles_and_records(gpu, tile, N1, N2, eps):
    """complex64 2-D NUFFT with clustered points: every gridding and
    degridding tile size (type 1 and type 2) against the FP32-ABI oracle;
    the MFMA quadrant gridder (tiles of 16 or 32) against the LDS-atomic one;
    the two record paths bit-identical, and type 2 unchanged by the type-1
    options.  32 x 48 modes give a grid of two tiles along x at tile 32."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(40 + tile + N1)
    M = 60000
    x, y = _clustered_xy(rng, M)
    c = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    fk = (rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))).astype(np.complex64)
    tx, ty, tc, tf = (torch.from_numpy(a).cuda() for a in (x, y, c, fk))
    out = {}
    for prerec, mfma in ((1, 1), (0, 1), (1, 0)):
        ctx = Context(tile_grid2d=tile, tile_degrid2d=tile, prerec=prerec, mfma2d=mfma)
        try:
            f1 = ctx.nufft2d1(tx, ty, tc, N1, N2, eps).cpu().numpy()
            c2 = ctx.nufft2d2(tx, ty, tf, eps).cpu().numpy()
        finally:
            ctx.close()
        out[prerec, mfma] = (f1, c2)
    assert np.array_equal(out[1, 1][0], out[0, 1][0])
    assert np.array_equal(out[1, 1][1], out[0, 1][1])
    assert np.array_equal(out[1, 1][1], out[1, 0][1])
    # FP32 MFMA sums against FP64 LDS sums of the same FP32 terms
    assert rel_l2(out[1, 1][0], out[1, 0][0]) <= 1e-6
    out = {1: out[1, 1]}
    ref1 = O.nufft2d1(x, y, c, N1, N2, as_float_abi(eps), single=True)
    ref2 = O.nufft2d2(x, y, fk, as_float_abi(eps), single=True)
    assert rel_l2(out[1][0], ref1) <= max(eps, _floor(True))
    assert rel_l2(out[1][1], ref2) <= max(eps, _floor(True))


@pytest.mark.parametrize("npix", [512, 1024, 2048])
@pytest.mark.parametrize("double", [False, True])
def test_row_fft_prefetch_bit_identical(gpu, monkeypatch, npix, double):
    """The row FFT that loads the next live row while transforming the
    current one (rows of at most 512 threads: 512- to 4096-point grids here)
    gives the same ms2dirty image, dirty2ms visibilities and 2-D NUFFT
    outputs, to rounding, as the one-row-at-a-time kernel (CG_FFT_ROWPF=0),
    with the row mask skipping rows no point reaches."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(90 + npix)
    nrow, fov = 20000, 2.0
    uvw = np.empty((nrow, 3))
    uvw[:, 0] = rng.normal(0, 2000, nrow)
    uvw[:, 1] = rng.normal(0, 300, nrow)  # few grid rows are reached
    uvw[:, 2] = rng.uniform(-50, 50, nrow)
    cdt = np.complex128 if double else np.complex64
    vis = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(cdt)
    dirty = (rng.standard_normal((npix, npix)) + 1j * rng.standard_normal((npix, npix))).astype(cdt)
    x, y = _clustered_xy(rng, nrow)
    fk = (rng.standard_normal((npix // 2, npix // 2)) + 1j * rng.standard_normal((npix // 2, npix // 2))).astype(cdt)
    c = vis
    if double:
        x, y = x.astype(np.float64), y.astype(np.float64)
    tu, tv, td = (torch.from_numpy(a).cuda() for a in (uvw, vis, dirty))
    tx, ty, tc, tf = (torch.from_numpy(a).cuda() for a in (x, y, c, fk))
    eps = 1e-9 if double else 1e-5
    out = {}
    for pf in ("1", "0"):
        monkeypatch.setenv("CG_FFT_ROWPF", pf)
        ctx = Context()
        try:
            out[pf] = (ctx.ms2dirty(tu, tv, None, npix, npix, fov, 1.0e9, eps).cpu().numpy(),
                       ctx.dirty2ms(tu, td, None, fov, 1.0e9, eps).cpu().numpy(),
                       ctx.nufft2d1(tx, ty, tc, npix // 2, npix // 2, eps).cpu().numpy(),
                       ctx.nufft2d2(tx, ty, tf, eps).cpu().numpy())
        finally:
            ctx.close()
    # the two kernels are separate instruction streams whose compiled sums may
    # round differently (and FP64 gridding sums in LDS atomics whose order
    # varies from run to run): they agree to rounding, where a wrongly placed
    # or skipped row would differ at O(1)
    for a, b in zip(out["1"], out["0"]):
        assert rel_l2(a, b) <= (1e-13 if double else 1e-6)
    assert np.abs(out["1"][0]).max() > 0 and np.abs(out["1"][1]).max() > 0


@pytest.mark.parametrize("nchan", [1, 3])
@pytest.mark.parametrize("uvw32", [False, True])
def test_ms2dirty_record_paths_identical(gpu, nchan, uvw32):
    """complex64 ms2dirty: records built in the key pass and gathered, or
    gathered straight from uvw / vis / wgt, give the same image bit for bit
    (one and several channels, FP64 and FP32 uvw), with and without weights."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(70 + nchan)
    nrow, nx, ny, fov = 5000, 96, 64, 3.0
    freqs = np.linspace(1.0e9, 1.3e9, nchan)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 300, (nrow, 2))
    uvw[:, 2] = rng.uniform(-80, 80, nrow)
    shape = (nrow,) if nchan == 1 else (nrow, nchan)
    vis = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    tu = torch.tensor(uvw.astype(np.float32) if uvw32 else uvw, device="cuda")
    tv, tw = torch.from_numpy(vis).cuda(), torch.from_numpy(wgt).cuda()
    f = freqs[0] if nchan == 1 else freqs
    out = {}
    for prerec in (1, 0):
        ctx = Context(prerec=prerec)
        try:
            out[prerec] = (ctx.ms2dirty(tu, tv, tw, nx, ny, fov, f, 1e-5).cpu().numpy(),
                           ctx.ms2dirty(tu, tv, None, nx, ny, fov, f, 1e-5).cpu().numpy())
        finally:
            ctx.close()
    assert np.array_equal(out[1][0], out[0][0])
    assert np.array_equal(out[1][1], out[0][1])


@pytest.mark.parametrize("spread,nchan,npix,band_max", [(150.0, 1, 1024, 1024), (2500.0, 1, 1024, 1024),
                                                       (250.0, 3, 1024, 1024), (9000.0, 1, 1024, 0),
                                                       (3500.0, 1, 2048, 2048), (3000.0, 2, 2048, 2048)])
def test_band_column_fft_matches_four_step(gpu, spread, nchan, npix, band_max):
    """complex64 images with several w-planes: the band-limited column pass
    (the rows points reach inside one band of at most 1024 rows, or 2048 when
    nf2 >= 4096) against the four-step pass (option band_fft = 0), both
    directions, and adjointness.  band_max 0: the band is too wide and both
    runs take the four-step pass.  1200 x 1024 (nf2 not a power of two)
    always takes the four-step pass."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(int(spread) + nchan + npix)
    nrow, fov = 60000, 2.0 * npix / 1024
    freqs = np.linspace(1.0e8, 1.15e8, nchan)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, spread, (nrow, 2))
    uvw[:, 2] = rng.uniform(-60, 60, nrow)
    shape = (nrow,) if nchan == 1 else (nrow, nchan)
    vis = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    tu = torch.tensor(uvw, device="cuda")
    tv, tw = torch.from_numpy(vis).cuda(), torch.from_numpy(wgt).cuda()
    f = freqs[0] if nchan == 1 else freqs
    shapes = [(npix, npix)] + ([(1200, 1024)] if npix == 1024 else [])
    for nx, ny in shapes:
        img = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
        ti = torch.from_numpy(img).cuda()
        out = {}
        for band in (1, 0):
            ctx = Context(band_fft=band)
            try:
                out[band] = (ctx.ms2dirty(tu, tv, tw, nx, ny, fov, f, 1e-5).cpu().numpy(),
                             ctx.dirty2ms(tu, ti, tw, fov, f, 1e-5).cpu().numpy())
                nf2, nf3 = ctx.last_plan()["nf2"], ctx.last_plan()["nf3"]
                bs, blen = ctx.last_band()
            finally:
                ctx.close()
            if band:
                # the planned band holds every row the points reach
                assert 0 < blen < nf2 and 0 <= bs < nf2
                if nx == npix and band_max:
                    assert blen <= band_max and nf2 >= 2 * band_max, (blen, nf2)
                elif nx == npix:
                    assert blen > 1024 and nf2 < 4096
        assert nf3 > 1
        assert rel_l2(out[1][0], out[0][0]) <= 1e-6
        assert rel_l2(out[1][1], out[0][1]) <= 1e-6
        lhs = np.vdot(out[1][1].astype(np.complex128).ravel(), vis.astype(np.complex128).ravel())
        rhs = np.vdot(img.astype(np.complex128).ravel(), out[1][0].astype(np.complex128).ravel())
        assert abs(lhs - rhs) <= 5e-5 * abs(lhs)


@pytest.mark.parametrize("single", [False, True])
def test_compact_keys_bit_identical(gpu, single):
    """Sort keys that number only the tiles of the u/v band rectangle give the
    same permutation: visibilities are bit-identical to the full numbering,
    images equal up to the order of the gridders' atomic sums (FP64 and
    complex64, dense gridding path active)."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(91)
    nrow, n, fov, f0 = 100000, 1024, 2.0, 1.0e8
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 1500, (nrow, 2))
    uvw[:, 2] = rng.uniform(-200, 200, nrow)
    cdt = np.complex64 if single else np.complex128
    vis = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(cdt)
    img = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(cdt)
    tu, tv, ti = (torch.from_numpy(a).cuda() for a in (uvw, vis, img))
    eps = 1e-5 if single else 1e-9
    out = {}
    for ck in (1, 0):
        ctx = Context(compact_keys=ck)
        try:
            out[ck] = (ctx.ms2dirty(tu, tv, None, n, n, fov, f0, eps).cpu().numpy(),
                       ctx.dirty2ms(tu, ti, None, fov, f0, eps).cpu().numpy())
            keys, _, _ = ctx.sorted_view()
            out[ck] += (int(keys.max()),)
        finally:
            ctx.close()
    assert out[1][2] < out[0][2]  # the compact numbering is in use
    assert rel_l2(out[1][0], out[0][0]) <= (1e-6 if single else 1e-13)
    assert np.array_equal(out[1][1], out[0][1])


def test_cfg1_plain_full_size(gpu):
    """The plain one-channel workload (131072 rows, 1024^2, fov 2 deg,
    freq 1e8, eps 1e-5, mixed precision) against the FP64 restatement on the
    whole image and the direct sum on sampled pixels/rows."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(1)
    M = 131072
    uvw = np.empty((M, 3))
    uvw[:, 0] = rng.normal(0, 1500, M)
    uvw[:, 1] = rng.normal(0, 1500, M)
    uvw[:, 2] = rng.uniform(-300, 300, M)
    vis = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    wgt = rng.uniform(0.5, 1.5, M).astype(np.float32)
    ctx = Context()
    tu, tv, tw = (torch.tensor(a, device="cuda") for a in (uvw, vis, wgt))
    d = ctx.ms2dirty(tu, tv, tw, 1024, 1024, 2.0, 1e8, 1e-5).cpu().numpy()
    img = (rng.standard_normal((1024, 1024))).astype(np.complex64)
    v = ctx.dirty2ms(tu, torch.tensor(img, device="cuda"), tw, 2.0, 1e8, 1e-5).cpu().numpy()
    ref = O.ms2dirty(uvw, 1e8, vis.astype("c16"), wgt.astype("f8"), 1024, 1024, 2.0, 1e-5)
    assert rel_l2(d, ref) <= 1e-5
    refv = O.dirty2ms(uvw, 1e8, img.astype("c16"), wgt.astype("f8"), 2.0, 1e-5)
    assert rel_l2(v, refv) <= 1e-5
    pix = np.random.default_rng(5).choice(1024 * 1024, 1000, replace=False).astype(np.int64)
    px = 2.0 * np.pi / 180 / 1024
    t = O.dft_ms2dirty(uvw, 1e8, vis.astype("c16"), wgt.astype("f8"), 1024, 1024, px, px,
                       pixels=pix)
    assert rel_l2(d.ravel()[pix], t) < 3e-5
    a1 = np.vdot(vis.astype("c16"), v.astype("c16")).real
    a2 = np.vdot(d.astype("c16"), img.astype("c16")).real
    assert abs(a1 - a2) <= 1e-5 * abs(a2)


def test_error_codes_do_not_abort(gpu):
    import cugridder_amd as cg
    L = cg.lib()
    uvw = np.zeros((4, 3))
    vis = np.ones(4, np.complex128)
    d = np.zeros((16, 16), np.complex128)
    assert L.ms2dirty_1(4, 0, 16, 1.0, 1e9, uvw, vis, d, 1e-5, 2.0, -1) == 2
    assert L.ms2dirty_1(4, 16, 16, 1.0, 1e9, uvw, vis, d, 1e-5, 1.0, -1) == 2
    # eps below the FP64 floor: clamped, ier = 1, still computed
    assert L.ms2dirty_1(4, 16, 16, 1.0, 1e9, uvw, vis, d, 1e-20, 2.0, -1) == 1
    assert np.isfinite(d).all()


def test_single_row_and_tiny_image(gpu):
    import cugridder_amd as cg
    k = wst_case(1, 16, 16, 5.0)
    d = cg.ms2dirty(k["uvw"], [k["f0"]], k["ms"], None, 16, 16, k["xpix"], k["ypix"], 0, 0, 1e-8,
                    True)
    ref = O.ms2dirty(k["uvw"], k["f0"], k["ms"], None, 16, 16, 5.0, 1e-8).real
    assert rel_l2(d, ref) <= 1e-8


@pytest.mark.parametrize("single", [False, True])
def test_row_mask_is_exact(gpu, single):
    """Skipping grid rows that no point reaches (zero before the forward FFT,
    unread after the inverse one) changes nothing: the degridding side is
    bit-identical; the gridding side differs only by the order of atomic
    accumulation.  The uv coverage spans a small fraction of the rows so
    that rows are skipped."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(11)
    nrow, nx, ny, fov, f0 = 20000, 512, 384, 2.0, 1e8
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 300, (nrow, 2))
    uvw[:, 2] = rng.uniform(-100, 100, nrow)
    cdt = torch.complex64 if single else torch.complex128
    rdt = torch.float32 if single else torch.float64
    tu = torch.tensor(uvw, device="cuda")
    tv = torch.tensor(rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow), dtype=cdt, device="cuda")
    tw = torch.tensor(rng.uniform(0.5, 1.5, nrow), dtype=rdt, device="cuda")
    img = torch.tensor(rng.standard_normal((nx, ny)), dtype=cdt, device="cuda")
    x = torch.tensor(rng.normal(0, 0.2, nrow), dtype=rdt, device="cuda")
    y = torch.tensor(rng.normal(0, 0.2, nrow), dtype=rdt, device="cuda")
    fk = torch.tensor(rng.standard_normal((192, 256)) + 1j * rng.standard_normal((192, 256)),
                      dtype=cdt, device="cuda")
    eps = 1e-5 if single else 1e-9
    out = {}
    for m in (0, 1):
        ctx = Context(row_mask=m)
        d = ctx.ms2dirty(tu, tv, tw, nx, ny, fov, f0, eps).cpu().numpy()
        v = ctx.dirty2ms(tu, img, tw, fov, f0, eps).cpu().numpy()
        f1 = ctx.nufft2d1(x, y, tv, 256, 192, eps).cpu().numpy()
        c2 = ctx.nufft2d2(x, y, fk, eps).cpu().numpy()
        out[m] = (d, v, f1, c2)
    tol = 1e-6 if single else 1e-13
    assert rel_l2(out[1][0], out[0][0]) < tol
    assert rel_l2(out[1][2], out[0][2]) < tol
    assert np.array_equal(out[1][1], out[0][1])
    assert np.array_equal(out[1][3], out[0][3])


def _mc_case(nrow, nchan, seed=21):
    rng = np.random.default_rng(seed)
    freqs = np.linspace(1.0e9, 1.6e9, nchan)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.uniform(-40, 40, (nrow, 2))
    uvw[:, 2] = rng.uniform(-5, 5, nrow)
    ms = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    wgt = rng.uniform(0.5, 1.5, (nrow, nchan))
    return freqs, uvw, ms, wgt


def _expand(uvw, freqs):
    """(row, channel) points as rows at the first frequency: uvw * f_c / f_0."""
    return (uvw[:, None, :] * (freqs / freqs[0])[None, :, None]).reshape(-1, 3)


@pytest.mark.parametrize("prec", ["double", "mixed"])
@pytest.mark.parametrize("nchan", [1, 3, 8])
def test_multichannel_vs_oracle(gpu, prec, nchan):
    """nchan frequencies: every (row, channel) is a visibility at uvw f_c/c
    and the image sums them (NIFTY semantics); against the oracle on the
    expanded rows, plus adjointness."""
    torch = gpu
    from cugridder_amd import Context
    nrow, nx, ny, fov = 400, 64, 96, 5.0
    freqs, uvw, ms, wgt = _mc_case(nrow, nchan)
    single = prec == "mixed"
    eps = 1e-5 if single else 1e-10
    cdt = torch.complex64 if single else torch.complex128
    rdt = torch.float32 if single else torch.float64
    tu = torch.tensor(uvw, device="cuda")
    tv = torch.tensor(ms, dtype=cdt, device="cuda")
    tw = torch.tensor(wgt, dtype=rdt, device="cuda")
    dirty = np.random.default_rng(5).standard_normal((nx, ny)) + 0j
    td = torch.tensor(dirty, dtype=cdt, device="cuda")
    ctx = Context()
    d = ctx.ms2dirty(tu, tv, tw, nx, ny, fov, freqs, eps).cpu().numpy()
    v = ctx.dirty2ms(tu, td, tw, fov, freqs, eps).cpu().numpy()
    assert v.shape == ((nrow,) if nchan == 1 else (nrow, nchan))
    f32 = (lambda a: np.asarray(a, np.complex64).astype(np.complex128)) if single else (lambda a: a)
    w32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if single else (lambda a: a)
    ue = _expand(uvw, freqs)
    ref_d = O.ms2dirty(ue, freqs[0], f32(ms).ravel(), w32(wgt).ravel(), nx, ny, fov, eps,
                       single=single)
    ref_v = O.dirty2ms(ue, freqs[0], f32(dirty), w32(wgt).ravel(), fov, eps, single=single)
    tol = 2e-6 if single else 1e-11
    assert rel_l2(d, ref_d) < tol
    assert rel_l2(v.ravel(), ref_v) < tol
    # <ms, D(dirty)> = <G(ms), dirty>
    lhs = np.vdot(f32(ms).ravel(), v.ravel())
    rhs = np.vdot(d, f32(dirty))
    assert abs(lhs - rhs) <= (5e-5 if single else 1e-11) * abs(rhs)


def test_multichannel_host_surface(gpu):
    """cursl with nchan > 1 runs the multi-channel host entry points; it
    agrees with the device entry points and returns (nrow, nchan)."""
    torch = gpu
    import cugridder_amd as cg
    from cugridder_amd import Context
    nrow, nx, ny, fov, nchan = 300, 48, 48, 4.0, 4
    freqs, uvw, ms, wgt = _mc_case(nrow, nchan, seed=3)
    xp = fov * np.pi / 180 / nx
    d = cg.ms2dirty(uvw, freqs, ms, wgt, nx, ny, xp, xp, 0, 0, 1e-9, True)
    ref = Context().ms2dirty(torch.tensor(uvw, device="cuda"), torch.tensor(ms, device="cuda"),
                             torch.tensor(wgt, device="cuda"), nx, ny, fov, freqs, 1e-9)
    assert rel_l2(d, ref.real.cpu().numpy()) < 1e-13
    img = np.random.default_rng(2).standard_normal((nx, ny))
    v = cg.dirty2ms(uvw, freqs, img, wgt, xp, xp, 0, 0, 1e-9, True)
    assert v.shape == (nrow, nchan)
    refv = Context().dirty2ms(torch.tensor(uvw, device="cuda"),
                              torch.tensor(img + 0j, device="cuda"),
                              torch.tensor(wgt, device="cuda"), fov, freqs, 1e-9).cpu().numpy()
    assert rel_l2(v, refv) < 1e-13


def test_multichannel_rejects_bad_frequency(gpu):
    from cugridder_amd import _lib
    L = _lib.lib()
    uvw = np.zeros((4, 3))
    ms = np.ones((4, 2), np.complex128)
    dirty = np.zeros((16, 16), np.complex128)
    ier = L.ms2dirty_mc_2(4, 2, 16, 16, 2.0, np.array([1e9, -1e9]), uvw, ms, None, dirty, 1e-6,
                          2.0, -1)
    assert ier == 2


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_calls(gpu, graph):
    """A batch of independent channels (mixed image sizes, counts and
    frequencies; autotuned; captured as one graph) gives the single-call
    results on every replay, with eager work on the stream between replays."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(8)
    shapes = [(64, 64), (128, 96), (96, 128), (256, 256), (64, 64), (48, 80)]
    nrows = [3000, 500, 20000, 8000, 2900, 0]
    ctx = Context()
    items = []
    for (nx, ny), m in zip(shapes, nrows):
        uvw = np.empty((m, 3))
        uvw[:, :2] = rng.normal(0, 300, (m, 2))
        uvw[:, 2] = rng.uniform(-50, 50, m)
        vis = (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64)
        items.append(dict(uvw=torch.tensor(uvw, device="cuda"),
                          vis=torch.tensor(vis, device="cuda"),
                          wgt=torch.tensor(rng.uniform(0.5, 1.5, m).astype(np.float32), device="cuda"),
                          dirty=torch.tensor((rng.standard_normal((nx, ny)) + 0j).astype(np.complex64),
                                             device="cuda"),
                          fov=float(rng.uniform(1.0, 4.0)), freq=float(rng.uniform(1e8, 3e8))))
    eps = 1e-5
    ref_d = [ctx.ms2dirty(it["uvw"], it["vis"], it["wgt"], *it["dirty"].shape, it["fov"], it["freq"],
                          eps).cpu().numpy() for it in items]
    ref_v = [ctx.dirty2ms(it["uvw"], it["dirty"], it["wgt"], it["fov"], it["freq"], eps).cpu().numpy()
             for it in items]
    # ms2dirty batch writes into copies of the images
    out_items = [dict(it, dirty=torch.zeros_like(it["dirty"])) for it in items]
    b1 = ctx.batch("ms2dirty", out_items, eps, graph=graph)
    info, is_graph = b1.info()
    assert is_graph == graph and info.shape == (len(items), 5)
    vis_items = [dict(it, vis=torch.zeros_like(it["vis"])) for it in items]
    b2 = ctx.batch("dirty2ms", vis_items, eps, graph=graph)
    for rep in range(3):
        for it in out_items:
            it["dirty"].fill_(7.0)
        for it in vis_items:
            it["vis"].fill_(-3.0)
        b1.run()
        b2.run()
        torch.cuda.synchronize()
        for k, it in enumerate(out_items):
            assert rel_l2(it["dirty"].cpu().numpy(), ref_d[k]) < 2e-6, (rep, k)
        for k, it in enumerate(vis_items):
            if nrows[k]:
                assert rel_l2(it["vis"].cpu().numpy(), ref_v[k]) < 2e-6, (rep, k)
    b1.close()
    b2.close()


@pytest.mark.parametrize("nx,ny", [(48, 8192), (40, 5625)])
def test_fp64_long_rows_vs_oracle(gpu, nx, ny):
    """complex128 images whose fast grid axis (nf1 = 2 ny) is beyond one
    workgroup's LDS (the radix-2 step around two half-length rows): both
    directions against the FP64 restatement, and adjointness."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(ny)
    nrow, fov, f0, eps = 3000, 1.0, 3e8, 1e-9
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 200.0, (nrow, 2))
    uvw[:, 2] = rng.uniform(-50, 50, nrow)
    vis = rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)
    img = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    ctx = Context()
    try:
        tu, tv, ti = (torch.from_numpy(a).cuda() for a in (uvw, vis, img))
        d = ctx.ms2dirty(tu, tv, None, nx, ny, fov, f0, eps)
        assert ctx.last_plan()["nf1"] > 8192
        v = ctx.dirty2ms(tu, ti, None, fov, f0, eps)
        d, v = d.cpu().numpy(), v.cpu().numpy()
    finally:
        ctx.close()
    ref_d = O.ms2dirty(uvw, f0, vis, None, nx, ny, fov, eps)
    ref_v = O.dirty2ms(uvw, f0, img, None, fov, eps)
    assert rel_l2(d, ref_d) <= eps and rel_l2(v, ref_v) <= eps
    lhs, rhs = np.vdot(vis, v), np.vdot(d, img)
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_fp64_nufft2d_long_rows_vs_oracle(gpu):
    """complex128 2-D NUFFT with 8192 modes along the fast axis (nf1 = 16384)."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(77)
    M, N1, N2, eps = 20000, 8192, 32, 1e-9
    x, y = (a.astype(np.float64) for a in _clustered_xy(rng, M))
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fk = rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))
    ctx = Context()
    try:
        tx, ty, tc, tf = (torch.from_numpy(a).cuda() for a in (x, y, c, fk))
        f1 = ctx.nufft2d1(tx, ty, tc, N1, N2, eps).cpu().numpy()
        assert ctx.last_plan()["nf1"] == 2 * N1
        c2 = ctx.nufft2d2(tx, ty, tf, eps).cpu().numpy()
    finally:
        ctx.close()
    assert rel_l2(f1, O.nufft2d1(x, y, c, N1, N2, eps)) <= eps
    assert rel_l2(c2, O.nufft2d2(x, y, fk, eps)) <= eps
