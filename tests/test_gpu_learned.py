"""GPU checks of the learned gridding kernel.

The HIP forward, backward (taps and deconvolution factors), training and
tabulation kernels are compared with the FP64 NumPy oracle
(oracle/learned.py); the trained kernel is then used by ms2dirty /
dirty2ms and compared with the FP64 direct DFT and with the ES kernel at
the same support (what the network gains), and the certification against
the ES certifier is shown to accept and to reject.
"""
import numpy as np
import pytest

from oracle import learned as LN
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L(gpu):
    from cugridder_amd import learned
    return learned


@pytest.fixture(scope="module")
def trained(L):
    """kw -> (params, losses), trained on the GPU from the ES kernel on
    2^20 seeded visibilities of a 256 x 256 patch."""
    samples = L.training_samples(n_vis=1 << 20, patch=256, seed=0)
    out = {}
    for kw in (4, 5, 6):
        print("training kw", kw, flush=True)
        out[kw] = L.train(kw, samples=samples)
    return out


def test_forward_matches_oracle(L):
    p = LN.init_params(3, out_scale=1.0)
    s = np.linspace(-1, 1, 1001).astype(np.float32)
    y = L.evaluate(p, s)
    y_ref, _ = LN.forward(p.astype(np.float64), s.astype(np.float64))
    assert np.max(np.abs(y - y_ref)) <= 2e-5 * (1 + np.max(np.abs(y_ref)))


@pytest.mark.parametrize("kw", [4, 5, 6, 7])
def test_gradient_matches_oracle(L, kw):
    p = LN.init_params(kw, out_scale=0.3)
    u, th = L.training_samples(n_vis=3000, seed=kw)
    loss, g = L.loss_and_grad(kw, p, u, th)
    p64 = p.astype(np.float64)
    loss_ref, g_ref = LN.loss_and_grad(p64, u, th, kw)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    assert np.linalg.norm(g - g_ref) <= 1e-3 * np.linalg.norm(g_ref)
    # the path through the correction factors is part of the gradient: the
    # oracle without it disagrees with the HIP result
    _, g_taps = LN.loss_and_grad(p64, u, th, kw, node_grad=False)
    assert np.linalg.norm(g - g_taps) > 1e-2 * np.linalg.norm(g_ref)


def test_es_start_is_the_es_kernel(L):
    """With the output layer at zero the learned kernel is the ES kernel:
    the GPU loss equals the oracle's ES functional."""
    kw = 6
    p = L.init_params(0)
    u, th = L.training_samples(n_vis=4096, seed=11)
    loss, _ = L.loss_and_grad(kw, p, u, th)
    F = LN.response_fn(LN.es_params_like(kw), u, th, kw)
    assert loss == pytest.approx(np.mean(np.abs(F - 1) ** 2), rel=1e-6)


@pytest.mark.parametrize("kw", [4, 5, 6])
def test_training_beats_es_at_same_support(L, trained, kw):
    p, losses = trained[kw]
    assert np.all(np.isfinite(losses))
    rng = np.random.default_rng(99)
    u, th = rng.uniform(0, 1, 20000), rng.uniform(-np.pi / 2, np.pi / 2, 20000)
    p64 = p.astype(np.float64)
    rms = np.sqrt(np.mean(np.abs(LN.response(p64, u, th, kw) - 1) ** 2))
    rms_es = np.sqrt(np.mean(np.abs(LN.response_fn(LN.es_params_like(kw), u, th, kw) - 1) ** 2))
    print("kw", kw, "rms learned %.3e ES %.3e" % (rms, rms_es))
    assert rms < 0.9 * rms_es
    # the loss the GPU reports for the last steps is the oracle's functional
    assert np.mean(losses[-100:]) == pytest.approx(rms ** 2, rel=0.2)


@pytest.mark.parametrize("kw", [5, 6])
def test_tabulation_matches_network(L, trained, kw):
    """The gridders' form exp(beta sqrt(1 - x^2) + q(s)) of the trained
    kernel against the network: q(s) close to y(s) where psi matters, and
    the gridding response of the tabulated kernel within a tenth of the
    network kernel's own error; the quadrature node values."""
    p64 = trained[kw][0].astype(np.float64)
    coef, qpsi = L.tabulate(kw, trained[kw][0])
    assert coef.shape == (16,)
    s = np.linspace(-1, 1, 4001)
    y, _ = LN.forward(p64, s)
    q = np.polynomial.polynomial.polyval(s, coef.astype(np.float64))
    logpsi = LN.prior(s, kw) + y
    wgt = np.exp(logpsi - logpsi.max())
    assert np.max(np.abs(q - y) * wgt) <= 5e-6

    def psi_tab(d):
        sd = LN.mlp_input(d, kw)
        v = np.exp(LN.prior(sd, kw) + np.polynomial.polynomial.polyval(sd, coef.astype(np.float64)))
        return np.where(np.abs(d) < 0.5 * kw, v, 0.0)

    rng = np.random.default_rng(7)
    u, th = rng.uniform(0, 1, 20000), rng.uniform(-np.pi / 2, np.pi / 2, 20000)
    F_net = LN.response(p64, u, th, kw)
    F_tab = LN.response_fn(psi_tab, u, th, kw)
    err_net = np.sqrt(np.mean(np.abs(F_net - 1) ** 2))
    assert np.sqrt(np.mean(np.abs(F_tab - F_net) ** 2)) <= 0.1 * err_net
    xq, _ = LN.quadrature(kw)
    assert np.allclose(qpsi, LN.psi(p64, xq, kw), rtol=3e-6, atol=0)


def _problem(L, patch, nvis, seed):
    import torch
    uvw, vis, fov, freq = L.validation_problem(patch=patch, nvis=nvis, seed=seed)
    dev = torch.device("cuda")
    return uvw, vis, fov, freq, torch.from_numpy(uvw).to(dev), torch.from_numpy(vis).to(dev)


def test_learned_imaging_vs_dft(L, trained):
    """ms2dirty / dirty2ms with the trained support-6 kernel against the FP64
    direct DFT: within 1e-5 and below the ES kernel of the same support."""
    import torch
    from cugridder_amd import Context
    kw, eps, n = 6, 1e-5, 128
    uvw, vis, fov, freq, uvw_t, vis_t = _problem(L, n, 4096, seed=3)
    pix = fov * np.pi / 180 / n
    ref = O.dft_ms2dirty(uvw, freq, vis, None, n, n, pix, pix).reshape(n, n)
    learned = Context()
    es = Context(use_learned=0)
    try:
        learned.set_kernel(kw, trained[kw][0], eps)
        d = learned.ms2dirty(uvw_t, vis_t, None, n, n, fov, freq, eps)
        assert learned.last_plan()["kw"] == kw
        d_es = es.ms2dirty(uvw_t, vis_t, None, n, n, fov, freq, eps)
        assert es.last_plan()["kw"] == kw
        torch.cuda.synchronize()
        err, err_es = rel_l2(d.cpu().numpy(), ref), rel_l2(d_es.cpu().numpy(), ref)
        print("ms2dirty vs DFT: learned %.3e ES %.3e" % (err, err_es))
        assert err <= eps and err < err_es
        rng = np.random.default_rng(4)
        img = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
        img_t = torch.from_numpy(img).to(uvw_t.device)
        v = learned.dirty2ms(uvw_t, img_t, None, fov, freq, eps).cpu().numpy()
        v_ref = O.dft_dirty2ms(uvw, freq, img.astype(np.complex128), None, pix, pix)
        v_es = es.dirty2ms(uvw_t, img_t, None, fov, freq, eps).cpu().numpy()
        e2, e2_es = rel_l2(v, v_ref), rel_l2(v_es, v_ref)
        print("dirty2ms vs DFT: learned %.3e ES %.3e" % (e2, e2_es))
        assert e2 <= eps and e2 < e2_es
        # adjointness <dirty2ms(img), vis> = <img, ms2dirty(vis)>
        lhs = np.vdot(v.astype(np.complex128), vis.astype(np.complex128))
        rhs = np.vdot(img.astype(np.complex128), d.cpu().numpy().astype(np.complex128))
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
        # removing the kernel goes back to the ES kernel of the call's eps
        learned.set_kernel(kw, None, 0)
        learned.ms2dirty(uvw_t, vis_t, None, n, n, fov, freq, 1e-6)
        assert learned.last_plan()["kw"] == 7
    finally:
        learned.close()
        es.close()


def test_learned_kernel_on_the_mfma_paths(L, trained):
    """A learned kernel in the MFMA gridder and degridder (every bin dense)
    takes per-tap polynomials fitted to its values; against the sparse
    kernels, which evaluate the learned exponent exactly, and against the
    FP64 direct DFT."""
    import torch
    from cugridder_amd import Context
    kw, eps, n = 6, 1e-5, 128
    uvw, vis, fov, freq, uvw_t, vis_t = _problem(L, n, 4096, seed=5)
    pix = fov * np.pi / 180 / n
    ref = O.dft_ms2dirty(uvw, freq, vis, None, n, n, pix, pix).reshape(n, n)
    rng = np.random.default_rng(6)
    img = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    img_t = torch.from_numpy(img).to(uvw_t.device)
    v_ref = O.dft_dirty2ms(uvw, freq, img.astype(np.complex128), None, pix, pix)
    out = {}
    for dense in (1, 0):
        ctx = Context(dense_min=dense, degrid_dense_min=dense)
        try:
            ctx.set_kernel(kw, trained[kw][0], eps)
            d = ctx.ms2dirty(uvw_t, vis_t, None, n, n, fov, freq, eps).cpu().numpy()
            assert ctx.last_plan()["kw"] == kw
            v = ctx.dirty2ms(uvw_t, img_t, None, fov, freq, eps).cpu().numpy()
            out[dense] = (d, v)
        finally:
            ctx.close()
    (d1, v1), (d0, v0) = out[1], out[0]
    print("MFMA vs sparse: %.3e %.3e; vs DFT %.3e %.3e"
          % (rel_l2(d1, d0), rel_l2(v1, v0), rel_l2(d1, ref), rel_l2(v1, v_ref)))
    assert rel_l2(d1, d0) <= 2e-6 and rel_l2(v1, v0) <= 2e-6
    assert rel_l2(d1, ref) <= eps and rel_l2(v1, v_ref) <= eps


def test_certification_accepts_and_rejects(L, trained):
    """The ES certifier (support 7, eps 1e-6) accepts the support-6 kernel at
    1e-5 (the learned workload's bar) and rejects supports 5 and 4 there
    (their per-axis error alone is above it)."""
    ok = L.certify(6, trained[6][0], 1e-5, patch=256, nvis=1 << 15)
    print("certify kw 6:", ok)
    assert ok["accepted"] and 0 < ok["ms2dirty"] <= 1e-5 and ok["dirty2ms"] <= 1e-5
    for kw in (5, 4):
        bad = L.certify(kw, trained[kw][0], 1e-5, patch=256, nvis=1 << 15)
        print("certify kw %d:" % kw, bad)
        assert not bad["accepted"]


def test_learned_kernel_only_at_its_sigma(L, trained):
    """A kernel trained and certified at sigma 2 is not used by a call at
    another oversampling factor: that call plans the ES kernel of its own
    sigma (support from the sigma != 2 rule of setup_conv_opts)."""
    import torch
    from cugridder_amd import Context
    uvw, vis, fov, freq, uvw_t, vis_t = _problem(L, 64, 2000, seed=8)
    ctx = Context()
    try:
        ctx.set_kernel(6, trained[6][0], 1e-5)
        ctx.ms2dirty(uvw_t, vis_t, None, 64, 64, fov, freq, 1e-5)
        assert ctx.last_plan()["kw"] == 6
        ctx.ms2dirty(uvw_t, vis_t, None, 64, 64, fov, freq, 1e-5, sigma=1.5)
        _, kw_es, _, _ = O.setup_conv(1e-5, 1.5, single=True)
        assert ctx.last_plan()["kw"] == kw_es and kw_es > 6
        torch.cuda.synchronize()
    finally:
        ctx.close()
