"""CPU tests: the Python binding rejects tensors the C ABI would misread
(wrong dtype, length, layout or device) before anything is launched.

The library reads and writes exactly the element counts it is given, so a
float32 cross-section array or a host tensor would otherwise be an
out-of-bounds device access.  No GPU is needed: every case raises before a
context is used.
"""
import numpy as np
import pytest
import torch

import mcgpu
from mcgpu.autograd import SlabProblem, slab_tally


class _NoLaunch:
    """A stand-in context: any attribute other than `device` is a launch."""

    def __init__(self, device=None):
        self.device = device

    def __getattr__(self, name):
        raise AssertionError(f"{name} was reached: validation did not stop the call")


def _bank(n, device="cpu"):
    return {"seed": torch.zeros(n, dtype=torch.int64, device=device),
            "x": torch.zeros(n, dtype=torch.float64, device=device),
            "mu": torch.ones(n, dtype=torch.float64, device=device),
            "wmc": torch.ones(n, dtype=torch.float64, device=device),
            "index": torch.zeros(n, dtype=torch.int32, device=device)}


def _bare_context(device=0):
    """A Context object without a library handle: only its checks can run."""
    c = mcgpu.Context.__new__(mcgpu.Context)
    c.device = device
    c._h = None
    return c


def test_autograd_rejects_float32_cross_sections():
    p = SlabProblem(_NoLaunch(), _bank(16), 0, 8, 0.125, 1e-3)
    sa = torch.rand(8)  # torch's default dtype is float32
    ss = torch.rand(8, dtype=torch.float64)
    with pytest.raises(TypeError, match="sig_a"):
        slab_tally(sa, ss, p)
    with pytest.raises(TypeError, match="sig_s"):
        slab_tally(ss, sa, p)


def test_autograd_rejects_wrong_length_and_hint():
    p = SlabProblem(_NoLaunch(), _bank(16), 0, 8, 0.125, 1e-3)
    ok = torch.rand(8, dtype=torch.float64)
    with pytest.raises(ValueError, match="7 elements, expected 8"):
        slab_tally(torch.rand(7, dtype=torch.float64), ok, p)
    with pytest.raises(ValueError, match="target_hint"):
        slab_tally(ok, ok, p, torch.rand(9, dtype=torch.float64))


def test_host_tensors_rejected_for_a_device_context():
    with pytest.raises(ValueError, match="expected cuda:0"):
        SlabProblem(_NoLaunch(device=0), _bank(4), 0, 8, 0.125, 1e-3)


def test_bank_dtypes_and_keys():
    b = _bank(4)
    b["index"] = b["index"].to(torch.int64)
    with pytest.raises(TypeError, match="index"):
        mcgpu.check_bank(b, "bank")
    b = _bank(4)
    b["wmc"] = torch.ones(5, dtype=torch.float64)
    with pytest.raises(ValueError, match="wmc"):
        mcgpu.check_bank(b, "bank")
    with pytest.raises(TypeError, match="keys"):
        mcgpu.check_bank({"x": b["x"]}, "bank")


def test_check_tensor_layout_and_type():
    t = torch.zeros(8, 2, dtype=torch.float64)[:, 0]
    with pytest.raises(ValueError, match="contiguous"):
        mcgpu.check_tensor(t, "t", torch.float64)
    with pytest.raises(TypeError, match="torch.Tensor"):
        mcgpu.check_tensor(np.zeros(8), "t", torch.float64)
    mcgpu.check_tensor(None, "t", torch.float64, optional=True)
    with pytest.raises(ValueError, match="required"):
        mcgpu.check_tensor(None, "t", torch.float64)


def test_context_methods_check_before_the_library():
    c = _bare_context(0)
    m = 8
    f64 = torch.zeros(m, dtype=torch.float64)
    with pytest.raises(ValueError, match="expected cuda:0"):
        c.simulate_device(_bank(4), f64, f64, f64, 0, m, 0.125, 1e-3)
    with pytest.raises(ValueError, match="expected cuda:0"):
        c.forward_cells(_bank(4), f64, f64, f64, 0, m, 0.125, 1e-3)
    with pytest.raises(ValueError, match="expected cuda:0"):
        c.adjoint_cells(_bank(4), f64, f64, 0, m, 0.125, 1e-3, f64, f64, f64)
    # a slab table that runs past the bank
    with pytest.raises(ValueError, match="slab 0"):
        mcgpu._check_slabs([(0, 8, 0.125, 1e-3, 0, 2, 4)], 5)
