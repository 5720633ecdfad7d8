"""FP64 NumPy restatement of the learned gridding kernel (TEST INFRASTRUCTURE ONLY).

The product library trains and tabulates the kernel with its own HIP kernels
(cugridder_amd/csrc/cg_learn.hip); this module is the CPU checker for them:
the same network, the same error functional and its gradient, in FP64.

Kernel: psi(d) = exp(beta (sqrt(1 - x^2) - 1) + y(s)), x = 2 d / kw,
s = 2 x^2 - 1, |d| < kw/2 (0 outside): the ES kernel's exponent (beta as
the reference picks it for sigma = 2, src/fourier/conv_interp_invoker.cu:
96-104) plus the network y = W3 . tanh(W2 tanh(W1 s + b1) + b2) + b3,
1 -> 32 -> 32 -> 1.  With W3 = b3 = 0 the kernel is the ES kernel.
Parameter vector (1153 FP32 values): W1[32] b1[32] W2[32][32] (row o holds
the weights into hidden unit o) b2[32] W3[32] b3.

Error functional: gridding a unit visibility at fractional grid offset u
with support kw and evaluating the corrected output at mode theta (radians
per grid cell) yields F(u, theta) times the exact value, with
  F = sum_t psi(d_t) e^{-i theta d_t} / psihat(theta),  d_t = ceil(u - kw/2) + t - u,
  psihat(theta) = sum_q g_q psi(x_q) cos(theta x_q)
the Gauss-Legendre rule the reference uses for the ES kernel's transform
(src/fourier/deconv.cu:12-65, half-line rule of p = floor(2 + 1.5 kw) nodes
doubled by symmetry).  loss = mean |F - 1|^2 over the samples.
"""
import math

import numpy as np

H = 32
W1, B1, W2, B2, W3, B3 = 0, 32, 64, 64 + 1024, 64 + 1024 + 32, 64 + 1024 + 64
NPARAMS = B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1], p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def init_params(seed=0, out_scale=1.0):
    """Glorot-style seeded initialisation; out_scale 0 gives the ES kernel."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    p[W1:B1] = rng.normal(0, 1.0, H)
    p[B1:W2] = rng.normal(0, 0.5, H)
    p[W2:B2] = rng.normal(0, 1.0 / math.sqrt(H), H * H)
    p[W3:B3] = out_scale * rng.normal(0, 1.0 / math.sqrt(H), H)
    return p


def es_beta(kw):
    """beta of the reference's ES kernel at sigma = 2
    (src/fourier/conv_interp_invoker.cu:96-104)."""
    return {2: 2.20, 3: 2.26, 4: 2.38}.get(kw, 2.30) * kw


def prior(s, kw):
    """ES exponent beta (sqrt(1 - x^2) - 1) as a function of s = 2 x^2 - 1."""
    s = np.asarray(s, np.float64)
    return es_beta(kw) * (np.sqrt(np.clip(0.5 * (1.0 - s), 0.0, None)) - 1.0)


def mlp_input(d, kw):
    x = 2.0 * np.asarray(d, np.float64) / kw
    return 2.0 * x * x - 1.0


def forward(p, s):
    """y(s) and the activations needed by backward()."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    s = np.asarray(s, np.float64).ravel()
    h1 = np.tanh(np.outer(s, w1) + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    return h2 @ w3 + b3, (s, h1, h2)


def backward(p, cache, dy):
    """Gradient of sum_e dy_e y(s_e) with respect to the parameter vector."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    s, h1, h2 = cache
    dy = np.asarray(dy, np.float64).ravel()
    g = np.zeros(NPARAMS)
    g[W3:B3] = dy @ h2
    g[B3] = dy.sum()
    dz2 = dy[:, None] * w3[None, :] * (1.0 - h2 * h2)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (dz2 @ w2) * (1.0 - h1 * h1)
    g[W1:B1] = dz1.T @ s
    g[B1:W2] = dz1.sum(0)
    return g


def psi(p, d, kw):
    d = np.asarray(d, np.float64)
    s = mlp_input(d, kw).ravel()
    y, _ = forward(p, s)
    return np.where(np.abs(d.ravel()) < 0.5 * kw, np.exp(prior(s, kw) + y), 0.0).reshape(d.shape)


def quadrature(kw):
    """Half-line Gauss-Legendre nodes x_q in [-kw/2, 0) and weights g_q such
    that psihat(theta) = 2 sum_q g_q psi(x_q) cos(theta x_q) for an even psi."""
    p = int(2 + 1.5 * kw)
    x, w = np.polynomial.legendre.leggauss(2 * p)
    return 0.5 * kw * x[:p], 0.5 * kw * w[:p]


def tap_offsets(u, kw):
    u = np.asarray(u, np.float64)
    return np.ceil(u - 0.5 * kw)[:, None] + np.arange(kw)[None, :] - u[:, None]


def psihat(p, theta, kw):
    xq, gq = quadrature(kw)
    return 2.0 * np.cos(np.outer(theta, xq)) @ (gq * psi(p, xq, kw))


def response(p, u, theta, kw):
    """F(u, theta) for paired samples."""
    d = tap_offsets(u, kw)
    v = psi(p, d, kw)
    a = (v * np.exp(-1j * theta[:, None] * d)).sum(1)
    return a / psihat(p, theta, kw)


def loss_and_grad(p, u, theta, kw, node_grad=True):
    """mean |F - 1|^2 and its gradient with respect to the parameters
    (node_grad False leaves out the path through psihat, for tests that show
    the deconvolution-factor backward pass is needed)."""
    u = np.asarray(u, np.float64)
    theta = np.asarray(theta, np.float64)
    n = u.size
    xq, gq = quadrature(kw)
    d = tap_offsets(u, kw)
    st = mlp_input(d, kw).ravel()
    yt, ct = forward(p, st)
    vt = np.where(np.abs(d.ravel()) < 0.5 * kw, np.exp(prior(st, kw) + yt), 0.0).reshape(d.shape)
    sq = mlp_input(xq, kw)
    yq, cq = forward(p, sq)
    vq = np.exp(prior(sq, kw) + yq)
    cq_mat = 2.0 * np.cos(np.outer(theta, xq)) * gq[None, :]
    ph = cq_mat @ vq
    ph_ = ph[:, None]
    ph_ = ph_[:, 0]
    e = (vt * np.exp(-1j * theta[:, None] * d)).sum(1) / ph_ - 1.0
    loss = float(np.mean(np.abs(e) ** 2))
    # dL/dpsi_t = 2 Re(conj(E) e^{-i th d_t}) / ph / n
    dpsi_t = 2.0 * np.real(np.conj(e)[:, None] * np.exp(-1j * theta[:, None] * d)) / ph_[:, None] / n
    # dL/dpsihat = -2 Re(conj(E) F) / ph / n; dpsihat/dpsi_q = cq_mat
    dph = -2.0 * np.real(np.conj(e) * (e + 1.0)) / ph_ / n
    dpsi_q = dph @ cq_mat
    g = backward(p, ct, (dpsi_t * vt).ravel())
    if node_grad:
        g = g + backward(p, cq, dpsi_q * vq)
    return loss, g


def es_params_like(kw, beta=None):
    """Reference ES kernel exp(beta (sqrt(1 - (2d/kw)^2) - 1)) on the same
    support, as a callable for comparisons (beta = 2.30 kw, the reference's
    sigma = 2 choice in src/fourier/conv_interp_invoker.cu:49-119)."""
    beta = 2.30 * kw if beta is None else beta

    def f(d):
        x = 2.0 * np.asarray(d, np.float64) / kw
        return np.where(np.abs(x) < 1, np.exp(beta * (np.sqrt(np.clip(1 - x * x, 0, None)) - 1.0)), 0.0)

    return f


def response_fn(fn, u, theta, kw):
    """F(u, theta) of an arbitrary even kernel callable fn(d)."""
    xq, gq = quadrature(kw)
    d = tap_offsets(u, kw)
    a = (fn(d) * np.exp(-1j * theta[:, None] * d)).sum(1)
    ph = 2.0 * np.cos(np.outer(theta, xq)) @ (gq * fn(xq))
    return a / ph


def adam(p, grad_fn, steps, lr, batches, b1=0.9, b2=0.999, eps=1e-8):
    """Reference Adam loop in FP64 (for prototyping and cross-checks)."""
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    hist = []
    for it in range(steps):
        L, g = grad_fn(p, batches[it % len(batches)])
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** (it + 1))
        vh = v / (1 - b2 ** (it + 1))
        p -= lr * mh / (np.sqrt(vh) + eps)
        hist.append(L)
    return p, np.array(hist)
