"""Learned gridding kernels: training data, GPU training, tabulation and
certification.

A learned kernel replaces the ES kernel exp(beta (sqrt(1 - x^2) - 1)),
x = 2d/kw (src/fourier/conv_interp_invoker.cu:49-119 picks its support and
beta) with psi(d) = exp(beta (sqrt(1 - x^2) - 1) + y(2 x^2 - 1)), y a
1 -> 32 -> 32 -> 1 tanh network with FP32 weights (1153 parameters; layout
in include/curafft.h) that starts at zero, i.e. at the ES kernel.  The same
network defines the correction factors through the Gauss-Legendre rule the
reference applies to the ES kernel (src/fourier/deconv.cu:80-94).

Training data is the gridding error the method itself produces.  Gridding a
visibility at fractional grid offset u (support kw), transforming and
correcting, and reading image pixel j of a P-pixel patch returns the exact
direct-DFT value times prod_axes F(u_axis, theta_j), with theta_j =
2 pi (j - P/2) / (sigma P) and
    F(u, theta) = sum_t psi(d_t) e^{-i theta d_t} / psihat(theta).
``training_samples`` draws the fractional offsets of seeded random
visibilities on the grid of a P x P patch and pairs each with a pixel of
that patch; the loss is mean |F - 1|^2 (the per-axis direct-DFT-vs-gridded
error; to first order the 2-D / 3-D error is the sum over axes).

Training runs on the GPU (cg_learn_train: HIP forward, backward and Adam
kernels).  ``certify`` then images a seeded P x P problem with the learned
kernel and with the ES kernel at support 7 (eps 1e-6, the certifier) and accepts
the kernel for eps only when both directions agree with the certifier within
eps (relative L2).  The ES kernel stays the fallback: a context only uses a
learned kernel that was installed with ``Context.set_kernel`` and whose
certified eps is no larger than the call's.
"""
import math

import numpy as np

from ._lib import check, lib

NPARAMS = 1153
H = 32
_W1, _B1, _W2, _B2, _W3, _B3 = 0, 32, 64, 64 + 1024, 64 + 1024 + 32, 64 + 1024 + 64
C_LIGHT = 299792458.0


def default_degree(kw):
    """Degree of the network's polynomial in s that the gridders evaluate."""
    return 15


def init_params(seed=0, out_scale=0.0):
    """Seeded random hidden layers; the output layer is scaled by out_scale
    (0: the network outputs 0 and the kernel is the ES kernel)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    p[_W1:_B1] = rng.normal(0, 1.0, H)
    p[_B1:_W2] = rng.normal(0, 0.5, H)
    p[_W2:_B2] = rng.normal(0, 1.0 / math.sqrt(H), H * H)
    p[_W3:_B3] = out_scale * rng.normal(0, 1.0 / math.sqrt(H), H)
    return p


def training_samples(n_vis=1 << 20, patch=256, sigma=2.0, seed=0):
    """(u, theta) pairs from n_vis seeded random visibilities on the grid of a
    patch x patch image: u is the fractional grid offset of a visibility
    along one axis (x for even, y for odd visibilities), theta the mode of a
    seeded random pixel of the patch."""
    rng = np.random.default_rng(seed)
    nf = int(round(sigma * patch))
    # visibilities spread over the whole uv-plane of the grid (cells)
    g = rng.uniform(-0.5 * nf, 0.5 * nf, size=(n_vis, 2))
    axis = np.arange(n_vis) & 1
    gc = g[np.arange(n_vis), axis]
    u = gc - np.floor(gc)
    j = rng.integers(0, patch, n_vis)
    theta = 2.0 * np.pi * (j - patch // 2) / (sigma * patch)
    return u.astype(np.float64), theta.astype(np.float64)


def train(kw, samples=None, batch=4096, schedule=((1e-3, 1500), (3e-4, 1500), (1e-4, 1500)),
          seed=0, params=None):
    """Trains the kernel of support kw on the GPU from the ES kernel (or from
    params); returns (params, losses).  The error functional is minimised
    with Adam over the schedule's (learning rate, steps) phases."""
    L = lib()
    p = init_params(seed) if params is None else np.array(params, np.float32)
    if samples is None:
        samples = training_samples(seed=seed)
    u, th = samples
    losses = []
    for lr, steps in schedule:
        h = np.zeros(steps)
        check(L.cg_learn_train(kw, 1, u, th, u.size, batch, steps, lr, p, h), "cg_learn_train")
        losses.append(h)
    return p, np.concatenate(losses) if losses else np.zeros(0)


def loss_and_grad(kw, params, u, theta):
    """Error functional and its gradient computed by the HIP kernels."""
    g = np.zeros(NPARAMS, np.float32)
    loss = np.zeros(1)
    check(lib().cg_learn_grad(kw, np.ascontiguousarray(params, np.float32),
                              np.ascontiguousarray(u, np.float64),
                              np.ascontiguousarray(theta, np.float64), len(u), g, loss),
          "cg_learn_grad")
    return float(loss[0]), g


def evaluate(params, s):
    s = np.ascontiguousarray(s, np.float32)
    y = np.zeros(s.size, np.float32)
    check(lib().cg_learn_eval(np.ascontiguousarray(params, np.float32), s, s.size, y), "cg_learn_eval")
    return y


def tabulate(kw, params, deg=None):
    """(coef[deg+1] of y as a polynomial in s, psi at the quadrature nodes)
    as the gridders use them."""
    deg = default_degree(kw) if deg is None else deg
    coef = np.zeros(deg + 1, np.float32)
    qpsi = np.zeros(int(2 + 1.5 * kw))
    check(lib().cg_learn_tabulate(kw, deg, np.ascontiguousarray(params, np.float32), coef, qpsi),
          "cg_learn_tabulate")
    return coef, qpsi


def validation_problem(patch=256, nvis=1 << 16, fov_deg=1.0, freq=1.0e9, seed=1):
    """Seeded w-stacking problem on a patch x patch image: uvw (float64,
    metres) spread over the image's uv-cells with |w| up to half the uv
    extent, complex64 visibilities with N(0, 1) parts."""
    rng = np.random.default_rng(seed)
    pix = fov_deg * np.pi / 180.0 / patch
    umax = 0.45 / pix * C_LIGHT / freq  # metres: just inside the grid's band
    uvw = np.empty((nvis, 3))
    uvw[:, 0] = rng.uniform(-umax, umax, nvis)
    uvw[:, 1] = rng.uniform(-umax, umax, nvis)
    uvw[:, 2] = rng.uniform(-0.5 * umax, 0.5 * umax, nvis)
    vis = (rng.standard_normal(nvis) + 1j * rng.standard_normal(nvis)).astype(np.complex64)
    return uvw, vis, fov_deg, freq


def certify(kw, params, eps, deg=None, patch=256, nvis=1 << 16, certifier_eps=1e-6, seed=1):
    """Relative L2 differences (ms2dirty, dirty2ms) of the learned kernel
    against the ES certifier on a seeded patch x patch problem, and whether both
    are within eps."""
    import torch
    from .context import Context
    deg = default_degree(kw) if deg is None else deg
    uvw_h, vis_h, fov, freq = validation_problem(patch, nvis, seed=seed)
    dev = torch.device("cuda")
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.from_numpy(vis_h).to(dev)
    certifier = Context(use_learned=0)
    cand = Context()
    try:
        cand.set_kernel(kw, params, eps, deg)
        d_ref = certifier.ms2dirty(uvw, vis, None, patch, patch, fov, freq, certifier_eps)
        d_new = cand.ms2dirty(uvw, vis, None, patch, patch, fov, freq, eps)
        if cand.last_plan()["kw"] != kw:
            raise RuntimeError("the learned kernel was not selected")
        rng = np.random.default_rng(seed + 1)
        img = torch.from_numpy((rng.standard_normal((patch, patch)) +
                                1j * rng.standard_normal((patch, patch))).astype(np.complex64)).to(dev)
        v_ref = certifier.dirty2ms(uvw, img, None, fov, freq, certifier_eps)
        v_new = cand.dirty2ms(uvw, img, None, fov, freq, eps)
        torch.cuda.synchronize()
        e1 = float(torch.linalg.norm(d_new - d_ref) / torch.linalg.norm(d_ref))
        e2 = float(torch.linalg.norm(v_new - v_ref) / torch.linalg.norm(v_ref))
    finally:
        certifier.close()
        cand.close()
    return {"ms2dirty": e1, "dirty2ms": e2, "accepted": bool(e1 <= eps and e2 <= eps)}
