"""The per-tap polynomials that the complex64 kernels may evaluate instead of
the exp2 form (cg_tap_poly, host code of the library): against the kernel
they stand for, evaluated directly in FP64 at offsets the fit never saw.

For tap t of a point whose first tap sits at offset d0 in [-kw/2, -kw/2 + 1),
sum_q c[t][q] x^q with x = 2 d0 + kw - 1 must equal psi(d0 + t) / e^beta,
psi the ES kernel (src/fourier/conv.cu:32-36) or the learned kernel.  The
interior taps are analytic on their interval and converge to FP32 rounding;
the two edge taps carry the square-root edge of the ES kernel and stay
within a few e^-beta (the kernel's own value at its edge).  Coefficients are
stored in FP32."""
import numpy as np
import pytest

from cugridder_amd import lib
from oracle import oracle as O


def _poly(kw, beta, c, deg, kq=None):
    out = np.zeros(kw * 12, np.float32)
    k = None if kq is None else np.ascontiguousarray(kq, np.float32)
    assert lib().cg_tap_poly(kw, beta, c, k, 0 if kq is None else len(kq) - 1, deg, out) == 0
    return out.reshape(kw, 12).astype(np.float64)


def _psi_norm(d, kw, beta, c, kq=None):
    u = c * d * d
    e = beta * (np.sqrt(np.maximum(1 - u, 0)) - 1)
    if kq is not None:
        e = e + np.polynomial.polynomial.polyval(2 * u - 1, kq)
    return np.where(np.abs(d) < kw / 2, np.exp(e), 0.0)


@pytest.mark.parametrize("eps", [1e-3, 1e-4, 1e-5, 1e-6])
def test_es_tap_polynomials(eps):
    _, kw, beta, c = O.setup_conv(eps, single=True)
    C = _poly(kw, beta, c, 10)
    rng = np.random.default_rng(kw)
    d0 = rng.uniform(-kw / 2, -kw / 2 + 1, 20000)
    x = 2 * d0 + kw - 1
    edge = 4 * np.exp(-beta)
    for t in range(kw):
        got = np.polynomial.polynomial.polyval(x, C[t])
        ref = _psi_norm(d0 + t, kw, beta, c)
        # interior taps: the FP32 rounding of the stored coefficients
        tol = edge if t in (0, kw - 1) else 5e-8
        assert np.abs(got - ref).max() <= tol, (kw, t, np.abs(got - ref).max())
    # a wrong tap order (coefficients of tap t used for tap t + 1) is far off
    assert np.abs(np.polynomial.polynomial.polyval(x, C[kw // 2]) -
                  _psi_norm(d0 + kw // 2 + 1, kw, beta, c)).max() > 0.1


def test_learned_tap_polynomials():
    """The learned kernel's exponent adds a polynomial in s = 2 c d^2 - 1."""
    kw, beta = 6, 13.8
    c = 4.0 / kw ** 2
    rng = np.random.default_rng(3)
    kq = rng.normal(0, 0.05, 16) / (1 + np.arange(16)) ** 2
    C = _poly(kw, beta, c, 10, kq)
    d0 = rng.uniform(-kw / 2, -kw / 2 + 1, 20000)
    x = 2 * d0 + kw - 1
    for t in range(1, kw - 1):
        got = np.polynomial.polynomial.polyval(x, C[t])
        ref = _psi_norm(d0 + t, kw, beta, c, kq.astype(np.float32).astype(np.float64))
        assert np.abs(got - ref).max() <= 1e-6, (t, np.abs(got - ref).max())


def test_tap_poly_rejects_bad_arguments():
    out = np.zeros(16 * 12, np.float32)
    assert lib().cg_tap_poly(0, 10.0, 0.1, None, 0, 10, out) == 2
    assert lib().cg_tap_poly(6, 10.0, 0.1, None, 0, 12, out) == 2
