"""Argument checks of the device-tensor API (cugridder_amd.context): host
tensors, wrong dtypes, strided views and short buffers are refused before
any pointer reaches the library (no GPU needed; nothing is launched)."""
import pytest
import torch

from cugridder_amd.context import _check, _check_imaging


def test_host_tensor_is_refused():
    with pytest.raises(TypeError, match="GPU"):
        _check(torch.zeros(4, 3, dtype=torch.float64), "uvw", (torch.float64,), shape=(None, 3))


def test_imaging_checks_refuse_before_device_checks_pass():
    uvw = torch.zeros(8, 3, dtype=torch.float64)
    vis = torch.zeros(8, dtype=torch.complex64)
    with pytest.raises(TypeError):
        _check_imaging(torch, uvw, vis, None, 1)
    with pytest.raises(TypeError, match="complex64 or complex128"):
        _check_imaging(torch, uvw, torch.zeros(8, dtype=torch.float32), None, 1)


def test_dtype_shape_and_size_errors():
    # the shape/size checks run after the device check, so exercise them on
    # a stand-in that reports is_cuda (meta tensors carry dtype and shape
    # but no storage)
    class Dev(torch.Tensor):
        @property
        def is_cuda(self):
            return True

    def dev(t):
        return t.as_subclass(Dev)

    with pytest.raises(TypeError, match="dtype"):
        _check(dev(torch.zeros(4, 3, dtype=torch.float32)), "uvw", (torch.float64,))
    with pytest.raises(ValueError, match="shape"):
        _check(dev(torch.zeros(4, 2, dtype=torch.float64)), "uvw", (torch.float64,), shape=(None, 3))
    with pytest.raises(ValueError, match="contiguous"):
        _check(dev(torch.zeros(3, 4, dtype=torch.float64).t()), "uvw", (torch.float64,))
    with pytest.raises(ValueError, match="elements"):
        _check(dev(torch.zeros(7, dtype=torch.complex64)), "vis", (torch.complex64,), numel=8)
    # complex64 data need FP32 weights; complex128 data need FP64 uvw
    uvw = dev(torch.zeros(8, 3, dtype=torch.float64))
    with pytest.raises(TypeError, match="wgt"):
        _check_imaging(torch, uvw, dev(torch.zeros(8, dtype=torch.complex64)),
                       dev(torch.zeros(8, dtype=torch.float64)), 1)
    with pytest.raises(TypeError, match="uvw"):
        _check_imaging(torch, dev(torch.zeros(8, 3, dtype=torch.float32)),
                       dev(torch.zeros(8, dtype=torch.complex128)), None, 1)
    with pytest.raises(ValueError, match="wgt"):
        _check_imaging(torch, uvw, dev(torch.zeros(8, dtype=torch.complex64)),
                       dev(torch.zeros(7, dtype=torch.float32)), 1)
    assert _check_imaging(torch, uvw, dev(torch.zeros(16, dtype=torch.complex64)),
                          dev(torch.zeros(16, dtype=torch.float32)), 2) == (0, 1, 0, 8)


def test_cursl_geometry_checks():
    """cursl refuses bad pixel sizes and image shapes before any call (the
    reference binding passes them on unchecked), and routes a pixel size
    along ny other than rad_pix_x * nx / ny, or do_wstacking=False, away
    from the reference entry points."""
    import numpy as np

    from cugridder_amd.cursl import _check_rows, _geometry
    with pytest.raises(ValueError, match="rad_pix_x"):
        _geometry(64, 32, 0.0, None, True)
    with pytest.raises(ValueError, match="rad_pix_y"):
        _geometry(64, 32, 1e-4, -1e-4, True)
    with pytest.raises(ValueError, match="positive integers"):
        _geometry(64, 0, 1e-4, None, True)
    assert _geometry(64, 32, 1e-4, 2e-4, True) == (2e-4, True)
    assert _geometry(64, 32, 1e-4, None, True) == (2e-4, True)
    assert _geometry(64, 32, 1e-4, 1e-4, True) == (1e-4, False)
    assert _geometry(64, 32, 1e-4, 2e-4, False) == (2e-4, False)
    with pytest.raises(ValueError, match="uvw"):
        _check_rows(np.zeros((5, 2)), None, "ms", 1)
    with pytest.raises(ValueError, match="ms must hold"):
        _check_rows(np.zeros((5, 3)), np.zeros(4), "ms", 1)
    _check_rows(np.zeros((5, 3)), np.zeros((5, 2)), "ms", 2)
