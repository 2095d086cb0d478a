"""Checks the MFMA lane layouts assumed by the convolution kernels
(similie_amd/csrc/mfma_layout.hpp) against a host matrix product."""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

LIB = Path(__file__).with_name("cpp") / "libmfma_probe.so"


def bf16_round(a):
    return torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).float().numpy()


@pytest.mark.parametrize("name,m,n,k", [("probe_mfma32", 32, 32, 16), ("probe_mfma16", 16, 16, 32)])
def test_mfma_lane_layout(name, m, n, k):
    lib = ctypes.CDLL(str(LIB))
    rng = np.random.default_rng(0)
    A = rng.standard_normal((m, k)).astype(np.float32)
    B = rng.standard_normal((k, n)).astype(np.float32)
    a = torch.from_numpy(A).cuda()
    b = torch.from_numpy(B).cuda()
    d = torch.full((m, n), np.nan, dtype=torch.float32, device="cuda")
    fn = getattr(lib, name)
    assert fn(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(d.data_ptr())) == 0
    ref = bf16_round(A).astype(np.float64) @ bf16_round(B).astype(np.float64)
    np.testing.assert_allclose(d.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)

