"""Kernel-level checks on the GPU through the C-ABI test hooks: the radix
sort and tile keys bit-exact against host FP64 / numpy stable sort, the plane
FFT against rocFFT (torch.fft), and the gridding and dirty2ms-prologue stages
against the FP64 oracle's intermediate grids."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2, wst_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,bits", [(0, 4), (1, 3), (100, 5), (4096, 12), (4097, 12),
                                    (100000, 20), (3000000, 18), (70000, 32)])
def test_radix_sort_is_stable_and_exact(gpu, m, bits):
    torch = gpu
    from cugridder_amd import Context
    ctx = Context()
    rng = np.random.default_rng(m + bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, hi, size=m, dtype=np.int64)
    if m > 10:
        keys[: m // 3] = keys[0]  # long runs of equal keys test stability
    tk = torch.tensor(keys.astype(np.uint32).view(np.int32), device="cuda")
    tv = torch.arange(m, dtype=torch.int32, device="cuda")
    ctx.sort_pairs(tk, tv, bits)
    torch.cuda.synchronize()
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(tv.cpu().numpy(), ref)
    assert np.array_equal(tk.cpu().numpy().view(np.uint32), keys[ref].astype(np.uint32))


def _host_keys(uvw, freq, plan, nf1, nf2, tx, ty, cs):
    """Sort key of each row, evaluated in FP64 on the host with the same
    operation order as the device (include/conv.h:24): the tile of the
    footprint start, then (cs > 0) the start cell inside the tile."""
    two_pi = 6.28318530717958648
    scale = 2.0 * np.pi * (freq / 299792458.0)

    def g(x, n):
        t = x - np.floor(x / two_pi) * two_pi
        t = t - np.where(t >= np.pi, two_pi, 0.0)
        return (t * 0.159154943091895336 + 0.5) * n

    kw = plan["kw"]
    x0 = np.ceil(g(uvw[:, 1] * scale * plan["pix_y"], nf1) - 0.5 * kw).astype(np.int64) % nf1
    y0 = np.ceil(g(uvw[:, 0] * scale * plan["pix_x"], nf2) - 0.5 * kw).astype(np.int64) % nf2
    tile = (y0 // ty) * (nf1 // tx) + x0 // tx
    if cs == 0:
        return tile
    assert cs == (tx * ty).bit_length() - 1
    return (tile << cs) | ((y0 % ty) * tx + x0 % tx)


@pytest.mark.parametrize("nrow,nx,ny,fov,eps,freq", [(5000, 256, 256, 10.0, 1e-6, 1e9),
                                                     (131072, 1024, 1024, 2.0, 1e-5, 1e8)])
def test_tile_keys_and_permutation_bit_exact(gpu, nrow, nx, ny, fov, eps, freq):
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(7)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 1500, (nrow, 2))
    uvw[:, 2] = rng.uniform(-300, 300, nrow)
    vis = np.ones(nrow, np.complex64)
    tu, tv = torch.tensor(uvw, device="cuda"), torch.tensor(vis, device="cuda")
    ctx = Context(stop_after=1, compact_keys=0)
    ctx.ms2dirty(tu, tv, None, nx, ny, fov, freq, eps)
    p = ctx.last_plan()
    keys, perm, ts = ctx.sorted_view()
    cs = ctx.key_shift
    # the start cell joins the key exactly when the dense gridding path runs
    assert (cs > 0) == (p["kw"] <= 8 and p["nf3"] <= 32)
    hk = _host_keys(uvw, freq, p, p["nf1"], p["nf2"], p["tile_x"], p["tile_y"], cs)
    ref = np.argsort(hk, kind="stable")
    assert np.array_equal(perm.cpu().numpy(), ref)
    assert np.array_equal(keys.cpu().numpy(), hk[ref])
    ntiles = (p["nf1"] // p["tile_x"]) * (p["nf2"] // p["tile_y"])
    assert np.array_equal(ts.cpu().numpy(), np.searchsorted(hk[ref] >> cs, np.arange(ntiles + 1)))
    # compact keys (tiles of the u / v bands numbered alone): the same
    # permutation and tile starts, with keys of fewer bits
    cc = Context(stop_after=1)
    cc.ms2dirty(tu, tv, None, nx, ny, fov, freq, eps)
    k2, perm2, ts2 = cc.sorted_view()
    assert np.array_equal(perm2.cpu().numpy(), ref)
    assert np.array_equal(ts2.cpu().numpy(), ts.cpu().numpy())
    assert int(k2.max()) <= int(keys.max())


@pytest.mark.parametrize("shape", [(1, 8, 8), (2, 30, 500), (3, 64, 2048), (1, 96, 16384),
                                   (2, 2000, 12), (1, 1250, 54), (4, 16, 5000),
                                   # power-of-two rows and split columns (compile-time plans)
                                   (1, 2048, 64), (1, 16384, 32), (1, 4096, 512), (2, 256, 1024),
                                   (1, 32768 // 64, 8192),
                                   # complex128 rows beyond one workgroup (radix-2 step
                                   # around two half-length rows)
                                   (1, 40, 11250), (2, 6, 16384)])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_fft2d_vs_rocfft(gpu, shape, dtype):
    torch = gpu
    from cugridder_amd import Context
    b, n2, n1 = shape
    dt = getattr(torch, dtype)
    ctx = Context()
    torch.manual_seed(0)
    x = torch.randn(b, n2, n1, dtype=dt, device="cuda")
    tol = 1e-6 if dt == torch.complex64 else 1e-14
    for d in (1, -1):
        y = x.clone()
        ctx.fft2d(y, d)
        ref = torch.fft.ifft2(x) * (n1 * n2) if d == 1 else torch.fft.fft2(x)
        assert (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item() < tol


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("case", [(3000, 64, 96, 10.0, 1e-6), (500, 100, 64, 20.0, 1e-3),
                                  (20000, 256, 256, 2.0, 1e-5)])
def test_gridding_stage_vs_oracle(gpu, single, case):
    torch = gpu
    from cugridder_amd import Context
    nrow, nx, ny, fov, eps = case
    if single and eps < 1e-5:
        eps = 1e-5
    k = wst_case(nrow, nx, ny, fov)
    cdt = torch.complex64 if single else torch.complex128
    ctx = Context(stop_after=1)
    tu = torch.tensor(k["uvw"], device="cuda")
    tv = torch.tensor(k["ms"], dtype=cdt, device="cuda")
    tw = torch.tensor(k["wgt"], dtype=torch.float32 if single else torch.float64, device="cuda")
    ctx.ms2dirty(tu, tv, tw, nx, ny, fov, k["f0"], eps)
    got = ctx.grid_view(cdt).cpu().numpy()
    ref = O.ms2dirty_grid(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, fov, eps, single=single)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < (3e-6 if single else 1e-13)


@pytest.mark.parametrize("case", [(3000, 64, 96, 10.0, 1e-6, 1, None),
                                  (20000, 256, 256, 2.0, 1e-5, 1, None),
                                  (131072, 1024, 1024, 2.0, 1e-5, 64, 1500.0),
                                  (20000, 128, 128, 5.0, 1e-3, 1, None),
                                  (50000, 200, 200, 3.0, 1e-4, 1, None),
                                  # crowded: ~33K points in each of a few bins,
                                  # so bins are cut into 4096-point units and
                                  # quads into several pieces
                                  (300000, 256, 256, 2.0, 1e-5, 1, 130.0)])
def test_dense_gridding_vs_oracle(gpu, case):
    """Bins holding at least dense_min points are gridded by MFMA products
    over quads of start cells (complex64, kw 4 to 7): the planes must match
    the oracle's spreading and the all-sparse path."""
    torch = gpu
    from cugridder_amd import Context
    nrow, nx, ny, fov, eps, dmin, sig = case
    if sig is not None:
        rng = np.random.default_rng(1)
        uvw = np.empty((nrow, 3))
        uvw[:, :2] = rng.normal(0, sig, (nrow, 2))
        uvw[:, 2] = rng.uniform(-300, 300, nrow)
        ms = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(np.complex64)
        wgt = rng.uniform(0.5, 1.5, nrow).astype(np.float32)
        f0 = 1e8
    else:
        k = wst_case(nrow, nx, ny, fov)
        uvw, ms, wgt, f0 = k["uvw"], k["ms"].astype(np.complex64), k["wgt"].astype(np.float32), k["f0"]
    tu = torch.tensor(uvw, device="cuda")
    tv = torch.tensor(ms, device="cuda")
    tw = torch.tensor(wgt, device="cuda")
    grids = {}
    for d in (0, dmin):
        ctx = Context(stop_after=1, dense_min=d)
        ctx.ms2dirty(tu, tv, tw, nx, ny, fov, f0, eps)
        grids[d] = ctx.grid_view(torch.complex64).cpu().numpy()
    p = ctx.last_plan()
    assert p["kw"] <= 8 and p["nf3"] <= 32, "case must be eligible for the dense path"
    ref = O.ms2dirty_grid(uvw, f0, ms.astype(np.complex128), wgt.astype(np.float64), nx, ny, fov,
                          eps, single=True)
    assert rel_l2(grids[dmin], ref) < 3e-6
    assert rel_l2(grids[dmin], grids[0]) < 3e-6


@pytest.mark.parametrize("case", [(131072, 1024, 1024, 2.0, 1e-5, 1500.0),
                                  (60000, 256, 256, 2.0, 1e-6, 300.0),
                                  (60000, 128, 160, 5.0, 1e-3, 300.0),
                                  (60000, 200, 200, 3.0, 1e-4, 300.0),
                                  # crowded: bins cut into units, quads into pieces
                                  (300000, 256, 256, 2.0, 1e-5, 130.0)])
def test_dense_degridding_vs_oracle(gpu, case):
    """dirty2ms with the tiles of many points interpolated by MFMA products
    over quads of start cells (complex64, kw 4 to 7, quads below the point
    threshold one lane per point) against the all-scalar degridder and the
    oracle."""
    torch = gpu
    from cugridder_amd import Context
    nrow, nx, ny, fov, eps, sig = case
    rng = np.random.default_rng(nrow + nx)
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, sig, (nrow, 2))
    uvw[:, 2] = rng.uniform(-300, 300, nrow)
    f0 = 1e8
    wgt = rng.uniform(0.5, 1.5, nrow).astype(np.float32)
    img = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
    tu, tw, ti = (torch.from_numpy(a).cuda() for a in (uvw, wgt, img))
    out = {}
    for d in (0, 1):
        ctx = Context(degrid_dense_min=d)
        try:
            out[d] = ctx.dirty2ms(tu, ti, tw, fov, f0, eps).cpu().numpy()
            p = ctx.last_plan()
        finally:
            ctx.close()
    assert p["kw"] <= 7 and p["nf3"] <= 32, "case must be eligible for the dense path"
    ref = O.dirty2ms(uvw, f0, img.astype(np.complex128), wgt.astype(np.float64), fov, eps, single=True)
    assert rel_l2(out[1], out[0]) < 2e-6
    assert rel_l2(out[1], ref) < max(eps, 2e-6)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("shape", [(48, 80), (100, 64), (160, 50), (128, 64)])
def test_dirty2ms_prologue_vs_oracle(gpu, single, shape):
    """The prologue writes the live columns (those holding image pixels) of
    every plane, with the column axis in the FFT's spectrum order; rows of
    frequencies outside the image rows are zero or, when the column FFT
    skips them, unwritten, so only image-row frequencies are compared."""
    torch = gpu
    from cugridder_amd import Context
    nx, ny = shape
    k = wst_case(200, nx, ny, 12.0)
    eps = 1e-5 if single else 1e-9
    cdt = torch.complex64 if single else torch.complex128
    ctx = Context(stop_after=1)
    img = torch.tensor(k["dirty"], dtype=cdt, device="cuda")
    ctx.dirty2ms(torch.tensor(k["uvw"], device="cuda"), img, None, 12.0, k["f0"], eps)
    got = ctx.grid_view(cdt).cpu().numpy()
    ref = O.dirty2ms_planes(k["uvw"], k["f0"], k["dirty"] + 0j, 12.0, eps, single=single)
    nf2, nf1 = ref.shape[1], ref.shape[2]
    live = np.r_[0:ny - ny // 2, nf1 - ny // 2:nf1]
    rows = Context.spectrum_rows(nf2)
    keep = (rows < nx - nx // 2) | (rows >= nf2 - nx // 2)
    g = got[:, keep][:, :, live]
    r = ref[:, rows[keep]][:, :, live]
    assert rel_l2(g, r) < (1e-6 if single else 1e-13)


@pytest.mark.parametrize("eps", [1e-5, 1e-6, 1e-4])
def test_tap_polynomial_kernel_values(gpu, eps):
    """The MFMA gridder and degridder with kernel values from per-tap
    polynomials (option tap_poly) against the exp2 form and the oracle."""
    torch = gpu
    from cugridder_amd import Context
    rng = np.random.default_rng(int(-np.log10(eps)))
    nrow, nx, ny, fov, f0 = 131072, 512, 512, 2.0, 1e8
    uvw = np.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0, 600, (nrow, 2))
    uvw[:, 2] = rng.uniform(-300, 300, nrow)
    vis = (rng.standard_normal(nrow) + 1j * rng.standard_normal(nrow)).astype(np.complex64)
    img = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))).astype(np.complex64)
    tu, tv, ti = (torch.from_numpy(a).cuda() for a in (uvw, vis, img))
    res = {}
    for tp in (0, 1):
        ctx = Context(tap_poly=tp, dense_min=1, degrid_dense_min=1)
        try:
            res[tp] = (ctx.ms2dirty(tu, tv, None, nx, ny, fov, f0, eps).cpu().numpy(),
                       ctx.dirty2ms(tu, ti, None, fov, f0, eps).cpu().numpy())
        finally:
            ctx.close()
    ref_d = O.ms2dirty(uvw, f0, vis.astype(np.complex128), None, nx, ny, fov, eps, single=True)
    ref_v = O.dirty2ms(uvw, f0, img.astype(np.complex128), None, fov, eps, single=True)
    for k, ref in ((0, ref_d), (1, ref_v)):
        assert rel_l2(res[1][k], res[0][k]) < max(0.2 * eps, 1e-6)
        assert rel_l2(res[1][k], ref) < max(eps, 2e-6)
