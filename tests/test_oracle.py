"""Pins the FP64 CPU oracle (oracle/cugrid_oracle.c) against independent
references: numpy FFT, numpy Gauss-Legendre, the reference's own
contrib/legendre_rule_fast.c (built into oracle/_ref), scipy quadrature,
brute-force sums written here in numpy, and the direct DFT."""
import warnings

import numpy as np
import pytest
from scipy import integrate

warnings.filterwarnings("ignore", category=integrate.IntegrationWarning)

from oracle import oracle as O
from tests.conftest import rel_l2, wst_case


def _smooth235(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def test_next235beven_bruteforce():
    for b in (1, 2, 4, 6):
        for n in range(0, 3000):
            r = O.next235beven(n, b)
            if n <= 2:
                assert r == 2
                continue
            assert r >= n and r % 2 == 0 and r % b == 0 and _smooth235(r)
            # minimality: no smaller admissible value
            for m in range(max(n, 2), r):
                assert not (m % 2 == 0 and m % b == 0 and _smooth235(m)), (n, b, m, r)


@pytest.mark.parametrize("n", [2, 3, 10, 11, 22, 41, 52])
def test_gauss_legendre(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14
    assert np.max(np.abs(w - wr)) < 1e-13
    # exact for polynomials of degree < 2n
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


@pytest.mark.parametrize("n", [10, 22, 41, 52])
def test_gauss_legendre_vs_reference_glr(n):
    ref = O.ref_glr(n)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    x, w = O.gauss_legendre(n)
    assert np.max(np.abs(x - ref[0])) < 1e-14
    assert np.max(np.abs(w - ref[1])) < 1e-14


def test_setup_conv_known_values():
    # kw = ceil(-log10(eps/10)), beta/kw = 2.30 (2.20/2.26/2.38 for kw 2/3/4)
    cases = {1e-5: (6, 13.8), 1e-6: (7, 16.1), 2e-1: (2, 4.4), 5e-3: (4, 9.52),
             5e-5: (6, 13.8), 5e-7: (8, 18.4), 5e-12: (13, 29.9), 1e-2: (3, 6.78)}
    for eps, (kw, beta) in cases.items():
        ier, k, b, c = O.setup_conv(eps)
        assert ier == 0 and k == kw and abs(b - beta) < 1e-12 and abs(c - 4.0 / kw ** 2) < 1e-15
    # clamping below the precision floor
    assert O.setup_conv(1e-20)[0] == 1 and O.setup_conv(1e-20)[1] == 16
    assert O.setup_conv(1e-9, single=True)[0] == 1


def test_set_nhg_w_invariants():
    for S, X, kw in [(1.5e-4, 629.0, 6), (3.4e-4, 31416.0, 7), (0.0038, 3.0, 11), (0.0, 0.0, 6),
                     (0.0, 5.0, 6), (1e-3, 0.0, 4)]:
        nf, h, g = O.set_nhg_w(S, X, kw)
        assert nf % 2 == 0 and nf >= 2 * kw
        assert abs(h * nf - 2 * np.pi) < 1e-12
        if X > 0 and S > 0:
            # w' = (w - c)/gamma stays inside (-pi, pi) and z' inside [-pi/sigma, pi/sigma]
            assert X / g <= np.pi
            # z' = (n-1 - o_c) gamma h stays inside [-pi/sigma, pi/sigma], with
            # equality when the safeguard max(S, 1/X) picks S
            assert S * g * h <= np.pi / 2.0 + 1e-12
            if S >= 1.0 / X:
                assert abs(S * g * h - np.pi / 2.0) < 1e-12


def _psi(x, kw, beta, c):
    return np.where(np.abs(x) < kw / 2, np.exp(beta * np.sqrt(np.maximum(1 - c * x * x, 0))), 0.0)


@pytest.mark.parametrize("eps", [1e-5, 1e-6, 5e-12])
def test_fseries_grid_vs_quadrature(eps):
    _, kw, beta, c = O.setup_conv(eps)
    nf = 64
    phi = O.fseries_grid(kw, beta, c, nf)
    for k in range(nf // 2 + 1):
        q, _ = integrate.quad(lambda x: _psi(x, kw, beta, c) * np.cos(2 * np.pi * k * x / nf),
                              -kw / 2, kw / 2, limit=200, epsabs=0, epsrel=1e-13)
        # the reference folds the fftshift sign (-1)^k into the factor; its
        # rule (p = 2 + 1.5 kw nodes on the half line) integrates psi, whose
        # derivatives are singular at +-kw/2, to ~1e-9 of phihat(0)
        assert abs(phi[k] - (-1) ** k * q) <= 1e-8 * abs(phi[0])


def test_fseries_at_vs_quadrature():
    _, kw, beta, c = O.setup_conv(1e-6)
    ks = np.array([0.0, 0.3, -0.7, 1.2, np.pi / 2])
    phi = O.fseries_at(kw, beta, c, ks)
    for k, p in zip(ks, phi):
        q, _ = integrate.quad(lambda x: _psi(x, kw, beta, c) * np.cos(k * x), -kw / 2, kw / 2,
                              limit=200, epsabs=0, epsrel=1e-13)
        assert abs(p - q) <= 1e-8 * abs(phi[0])


@pytest.mark.parametrize("n", [2, 6, 12, 30, 500, 2048, 2000, 1250])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(O.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel_l2(O.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-14


def test_fft2d_vs_numpy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 30, 64)) + 1j * rng.standard_normal((3, 30, 64))
    assert rel_l2(O.fft2d(x, -1), np.fft.fft2(x)) < 1e-14
    assert rel_l2(O.fft2d(x, +1), np.fft.ifft2(x) * 30 * 64) < 1e-14


def test_spread3d_vs_output_driven_bruteforce():
    """Input-driven periodic spreading equals the output-driven sum of
    conv_3d_outputdriven_t1 (each cell sums psi of the min-image distance)."""
    rng = np.random.default_rng(3)
    nf1, nf2, nf3, M = 16, 12, 10, 7
    _, kw, beta, c = O.setup_conv(1e-4)
    x = rng.uniform(-np.pi, np.pi, M)
    y = rng.uniform(-7, 7, M)          # outside [-pi, pi): exercises the fold
    z = rng.uniform(-np.pi, np.pi, M)
    cv = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    fw = O.spread3d(x, y, z, cv, nf1, nf2, nf3, kw, beta, c)

    def g(v, n):
        t = v - np.floor(v / (2 * np.pi)) * 2 * np.pi
        t = t - (t >= np.pi) * 2 * np.pi
        return (t / (2 * np.pi) + 0.5) * n

    def dist(b, gg, n):
        d = np.abs(b - gg)
        return np.where(d > n / 2, np.abs(n - d), d)

    ref = np.zeros((nf3, nf2, nf1), complex)
    for j in range(M):
        d1 = dist(np.arange(nf1), g(x[j], nf1), nf1)
        d2 = dist(np.arange(nf2), g(y[j], nf2), nf2)
        d3 = dist(np.arange(nf3), g(z[j], nf3), nf3)
        k1, k2, k3 = (_psi(d, kw, beta, c) for d in (d1, d2, d3))
        ref += cv[j] * k3[:, None, None] * k2[None, :, None] * k1[None, None, :]
    assert rel_l2(fw, ref) < 1e-14
    # interpolation is the adjoint of spreading
    u = rng.standard_normal((nf3, nf2, nf1)) + 1j * rng.standard_normal((nf3, nf2, nf1))
    ci = O.interp3d(x, y, z, u, kw, beta, c)
    assert abs(np.vdot(fw, u) - np.vdot(cv, ci)) < 1e-12 * abs(np.vdot(fw, u))


def _numpy_dft(uvw, f0, ms, wgt, nx, ny, xpix, ypix, sign=-1):
    """python/curagridder/test_wst.py:15-32 (complex result, weights applied)."""
    c = 299792458.0
    x, y = np.meshgrid(*[-ss / 2 + np.arange(ss) for ss in (nx, ny)], indexing="ij")
    x = x * xpix
    y = y * ypix
    eps = x ** 2 + y ** 2
    nm1 = -eps / (np.sqrt(1.0 - eps) + 1.0)
    res = np.zeros((nx, ny), complex)
    w = np.ones(len(ms)) if wgt is None else wgt
    for r in range(len(ms)):
        ph = f0 / c * (x * uvw[r, 0] + y * uvw[r, 1] + sign * uvw[r, 2] * nm1)
        res += ms[r] * w[r] * np.exp(2j * np.pi * ph)
    return res / (nm1 + 1)


def test_direct_dft_vs_numpy_formula():
    k = wst_case(27, 16, 24, 10.0)
    ref = _numpy_dft(k["uvw"], k["f0"], k["ms"], k["wgt"], 16, 24, k["xpix"], k["ypix"])
    got = O.dft_ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], 16, 24, k["xpix"], k["ypix"])
    assert rel_l2(got, ref) < 1e-13
    # the adjoint DFT is the conjugate transpose
    d = k["dirty"] + 0j
    v = O.dft_dirty2ms(k["uvw"], k["f0"], d, k["wgt"], k["xpix"], k["ypix"])
    assert abs(np.vdot(k["ms"], v) - np.vdot(got.reshape(16, 24), d)) < 1e-12 * np.abs(np.vdot(k["ms"], v))


@pytest.mark.parametrize("nrow,nx,ny,fov,eps", [(100, 32, 64, 10, 1e-10), (27, 16, 64, 20, 5e-7),
                                                (200, 64, 48, 1, 1e-5), (2, 32, 32, 5, 5e-3)])
@pytest.mark.parametrize("use_wgt", [False, True])
def test_restatement_vs_dft(nrow, nx, ny, fov, eps, use_wgt):
    k = wst_case(nrow, nx, ny, fov, use_wgt=use_wgt)
    # The method's own error against the exact sum (measured for the reference
    # algorithm restated in FP64): ms2dirty ~1.5-3 eps; dirty2ms on a few
    # hundred rows up to ~70 eps (small-sample L2 over few visibilities).
    d = O.ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, fov, eps)
    t = O.dft_ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, k["xpix"], k["ypix"])
    assert rel_l2(d, t) < 5 * eps
    dd = k["dirty"] + 0j
    v = O.dirty2ms(k["uvw"], k["f0"], dd, k["wgt"], fov, eps)
    tv = O.dft_dirty2ms(k["uvw"], k["f0"], dd, k["wgt"], k["xpix"], k["ypix"])
    assert rel_l2(v, tv) < 100 * eps
    # adjointness (python/curagridder/test_wst.py:84-89)
    assert abs(np.vdot(k["ms"], v).real - np.vdot(d, dd).real) <= 1e-12 * abs(np.vdot(d, dd).real)


def test_planted_sign_mistake_is_caught():
    """A wrong sign convention in the checker must be detected."""
    k = wst_case(100, 32, 32, 10)
    d = O.ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], 32, 32, 10, 1e-8)
    t_wrong = O.dft_ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], 32, 32, k["xpix"], k["ypix"], sign=+1)
    assert rel_l2(d, t_wrong) > 1e-3


@pytest.mark.parametrize("eps", [1e-4, 1e-9])
def test_nufft2d_vs_dft(eps):
    rng = np.random.default_rng(5)
    M, N1, N2 = 300, 24, 16
    x = rng.uniform(-np.pi, np.pi, M)
    y = rng.uniform(-np.pi, np.pi, M)
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    f = O.nufft2d1(x, y, c, N1, N2, eps)
    ref = O.dft2d1(x, y, c, N1, N2).reshape(N2, N1)
    # numpy brute force of test/nufft_2d_1_test.cu:276-283
    k1 = np.arange(N1) - N1 // 2
    k2 = np.arange(N2) - N2 // 2
    brute = np.einsum("j,ij,kj->ki", c, np.exp(1j * np.outer(k1, x)), np.exp(1j * np.outer(k2, y)))
    assert rel_l2(ref, brute) < 1e-13
    assert rel_l2(f, ref) < 10 * eps
    fk = rng.standard_normal((N2, N1)) + 1j * rng.standard_normal((N2, N1))
    cc = O.nufft2d2(x, y, fk, eps)
    assert rel_l2(cc, O.dft2d2(x, y, fk)) < 10 * eps


def test_sparse_image_dft_equals_dense_dft():
    """The sparse-image direct sum (used for 8192^2 images) equals the dense
    one on an image that is zero except at the listed pixels."""
    k = wst_case(40, 24, 32, 10.0)
    rng = np.random.default_rng(3)
    pix = rng.choice(24 * 32, 50, replace=False).astype(np.int64)
    val = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    img = np.zeros(24 * 32, np.complex128)
    img[pix] = val
    img = img.reshape(24, 32)
    for sign in (-1, 1):
        dense = O.dft_dirty2ms(k["uvw"], k["f0"], img, k["wgt"], k["xpix"], k["ypix"], sign=sign)
        rows = np.array([0, 7, 39], np.int64)
        sparse = O.dft_dirty2ms_sparse(k["uvw"], k["f0"], 24, 32, pix, val, k["wgt"], k["xpix"],
                                       k["ypix"], sign=sign, rows=rows)
        assert rel_l2(sparse, dense[rows]) < 1e-13


@pytest.mark.parametrize("nrow,nx,ny,fov,eps", [(60, 16, 24, 10.0, 1e-9), (200, 32, 20, 30.0, 1e-5)])
def test_plain_2d_restatement_vs_numpy_sum(nrow, nx, ny, fov, eps):
    """NIFTY's do_wstacking=False image, sum_j wgt_j vis_j e^{2 pi i f/c (u
    l + v m)} (w ignored, no 1/n), written here in numpy, against the
    oracle's 2-D NUFFT route (axis assignment u -> rows, v -> columns, and
    the sign); its adjoint against the conjugate transpose."""
    k = wst_case(nrow, nx, ny, fov)
    k["uvw"] *= 40.0  # phases well beyond a radian at the image edge
    l = (np.arange(nx) - nx // 2) * k["xpix"]
    m = (np.arange(ny) - ny // 2) * k["ypix"]
    s = 2 * np.pi * k["f0"] / 299792458.0
    E = np.exp(1j * s * (k["uvw"][:, 0][:, None, None] * l[None, :, None] +
                         k["uvw"][:, 1][:, None, None] * m[None, None, :]))
    c = k["ms"] * k["wgt"]
    brute = np.einsum("j,jab->ab", c, E)
    got = O.ms2dirty_2d(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, k["xpix"], k["ypix"], eps)
    assert rel_l2(got, brute) < 10 * eps
    # a transposed image (u and v swapped) is far off
    swapped = np.einsum("j,jab->ab", c, np.exp(1j * s * (
        k["uvw"][:, 1][:, None, None] * l[None, :, None] +
        k["uvw"][:, 0][:, None, None] * m[None, None, :])))
    assert rel_l2(got, swapped) > 1e-2
    d = k["dirty"] + 0.5j
    v = O.dirty2ms_2d(k["uvw"], k["f0"], d, k["wgt"], k["xpix"], k["ypix"], eps)
    vb = k["wgt"] * np.einsum("jab,ab->j", np.conj(E), d)
    assert rel_l2(v, vb) < 10 * eps


def test_pix_y_override_vs_dft():
    """The W-stacking restatement planned with a pixel size along ny other
    than fov / ny agrees with the direct sum at that pixel size and not at
    the reference's derived one."""
    nx, ny, fov, eps = 32, 48, 10.0, 1e-9
    k = wst_case(150, nx, ny, fov)
    xp = k["xpix"]
    yp = xp  # square pixels on a non-square image
    d = O.ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, fov, eps, pix_y=yp)
    t = O.dft_ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, xp, yp)
    assert rel_l2(d, t) < 5 * eps
    t_ref = O.dft_ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, xp, fov * np.pi / 180 / ny)
    assert rel_l2(d, t_ref) > 1e-3
    dd = k["dirty"] + 0j
    v = O.dirty2ms(k["uvw"], k["f0"], dd, k["wgt"], fov, eps, pix_y=yp)
    tv = O.dft_dirty2ms(k["uvw"], k["f0"], dd, k["wgt"], xp, yp)
    assert rel_l2(v, tv) < 100 * eps
    # the override is reset after the call
    d0 = O.ms2dirty(k["uvw"], k["f0"], k["ms"], k["wgt"], nx, ny, fov, eps)
    assert rel_l2(d0, t_ref) < 5 * eps
