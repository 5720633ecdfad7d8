"""Pins the learned-kernel oracle (oracle/learned.py) on the CPU.

The oracle is the checker of the HIP training kernels, so each of its parts
is checked against something outside its own formulas: the gradient against
central finite differences, the quadrature against numpy's Gauss-Legendre
rule and exact integrals, and the error functional F(u, theta) against a
brute-force 1-D gridding + FFT + correction of one point.
"""
import numpy as np
import pytest

from oracle import learned as LN


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    for kw in (4, 6):
        p = LN.init_params(3).astype(np.float64)
        u = rng.uniform(0, 8, 96)
        th = rng.uniform(-np.pi / 2, np.pi / 2, 96)
        _, g = LN.loss_and_grad(p, u, th, kw)
        for k in (0, 7, 40, 63, 100, 600, 1087, 1090, 1125, 1151):
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += e
            pm[k] -= e
            fd = (LN.loss_and_grad(pp, u, th, kw)[0] - LN.loss_and_grad(pm, u, th, kw)[0]) / (2 * e)
            assert g[k] == pytest.approx(fd, rel=2e-5, abs=1e-9), (kw, k)


def test_output_bias_has_no_effect():
    """F is a ratio of linear functionals of psi, so scaling psi (the output
    bias b3) changes nothing: its gradient vanishes."""
    rng = np.random.default_rng(2)
    p = LN.init_params(5).astype(np.float64)
    u, th = rng.uniform(0, 1, 64), rng.uniform(-1.5, 1.5, 64)
    l0, g = LN.loss_and_grad(p, u, th, 5)
    p[LN.B3] += 0.7
    l1, _ = LN.loss_and_grad(p, u, th, 5)
    assert abs(g[LN.B3]) < 1e-12 and l1 == pytest.approx(l0, rel=1e-12)


@pytest.mark.parametrize("kw", [4, 5, 6, 7])
def test_quadrature_rule(kw):
    xq, gq = LN.quadrature(kw)
    x, w = np.polynomial.legendre.leggauss(2 * len(xq))
    assert np.allclose(xq, 0.5 * kw * x[:len(xq)]) and np.all(xq < 0)
    # exact for even polynomials of degree < 4p: 2 sum g x^2k = kw^(2k+1)/(2^(2k)(2k+1))
    for k in range(4):
        exact = 2 * (0.5 * kw) ** (2 * k + 1) / (2 * k + 1)
        assert 2 * np.sum(gq * xq ** (2 * k)) == pytest.approx(exact, rel=1e-13)


@pytest.mark.parametrize("kw", [4, 6])
def test_error_functional_is_brute_force_gridding_error(kw):
    """One point at grid position g on a periodic grid of nf cells, spread with
    psi, FFT'd and divided by psihat: mode k holds e^{-i theta g} F(u, theta),
    theta = 2 pi k / nf, u = g - floor(g)."""
    rng = np.random.default_rng(4)
    p = LN.init_params(7).astype(np.float64)
    nf = 64
    for g in rng.uniform(10, 50, 5):
        grid = np.zeros(nf, complex)
        i0 = int(np.ceil(g - 0.5 * kw))
        for t in range(kw):
            grid[(i0 + t) % nf] += LN.psi(p, np.array([i0 + t - g]), kw)[0]
        spec = np.fft.fft(grid)  # sum_i grid[i] e^{-2 pi i k i / nf}
        k = np.arange(-nf // 4, nf // 4)
        theta = 2 * np.pi * k / nf
        got = spec[k % nf] / LN.psihat(p, theta, kw) * np.exp(1j * theta * g)
        want = LN.response(p, np.full(k.size, g - np.floor(g)), theta, kw)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
        # and the planted mistake (opposite phase convention) is caught
        bad = spec[k % nf] / LN.psihat(p, theta, kw) * np.exp(-1j * theta * g)
        assert not np.allclose(bad, want, rtol=1e-6)


def test_es_kernel_error_matches_its_published_accuracy():
    """With psi the reference's ES kernel (beta = 2.30 kw) the functional gives
    the accuracy the reference's support rule promises: kw = ceil(-log10(eps/10))
    reaches about eps / 10 .. eps in the rms."""
    rng = np.random.default_rng(5)
    u, th = rng.uniform(0, 1, 20000), rng.uniform(-np.pi / 2, np.pi / 2, 20000)
    for kw, eps in ((5, 1e-4), (6, 1e-5), (7, 1e-6)):
        F = LN.response_fn(LN.es_params_like(kw), u, th, kw)
        rms = np.sqrt(np.mean(np.abs(F - 1) ** 2))
        assert eps / 100 < rms < eps, (kw, rms)
