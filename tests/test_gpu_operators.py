"""GPU parity of operators._laplacian (operators.py:5-11) through wt_laplacian.

* Bitwise equal to the golden vectors written by the reference's own
  ``_laplacian`` (tests/golden/make_golden.py) and to the oracle's fp32
  restatement (pinned there by the same vectors, Dirichlet eigenvalues and
  quadratic fields) at h = 1 and at the reference example's h.
* Backward: the symmetric stencil with zero padding is self-adjoint, so the
  gradient equals the forward on the output gradient (bitwise, same kernel)
  and torch autograd through the restated reference op within 1e-6.
* float64: the reference's conv2d on the device within 1e-15 of the
  restated op on the CPU.
"""
import os

import numpy as np
import pytest
import torch

import oracle.c_oracle as co
from oracle import ref_torch
from tests.helpers import rel
from wavetorch_amd.operators import _laplacian

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_laplacian_golden_bitwise(cuda, gold):
    y = torch.from_numpy(gold["lap_in"]).to(cuda)
    out = _laplacian(y, torch.tensor(1.0))
    assert np.array_equal(out.cpu().numpy(), gold["lap_out"])


@pytest.mark.parametrize("h", [1.0, 1.4283556979968262, 0.37])
def test_laplacian_matches_oracle_bitwise(cuda, h):
    x = np.random.default_rng(7).standard_normal((3, 37, 41)).astype(np.float32)
    out = _laplacian(torch.from_numpy(x).to(cuda), h)
    assert np.array_equal(out.cpu().numpy(), co.laplacian(x, np.float32(h)))


def test_laplacian_backward(cuda):
    rng = np.random.default_rng(8)
    x = torch.from_numpy(rng.standard_normal((2, 23, 30)).astype(np.float32))
    w = torch.from_numpy(rng.standard_normal((2, 23, 30)).astype(np.float32))
    h = 1.25
    y = x.to(cuda).requires_grad_(True)
    (_laplacian(y, h) * w.to(cuda)).sum().backward()
    assert torch.equal(y.grad, _laplacian(w.to(cuda), h))
    xr = x.clone().requires_grad_(True)
    (ref_torch.laplacian(xr, torch.tensor(h)) * w).sum().backward()
    assert rel(y.grad.cpu().numpy(), xr.grad.numpy()) < 1e-6


def test_laplacian_float64(cuda):
    x = torch.from_numpy(np.random.default_rng(9).standard_normal((2, 19, 26)))
    h = 1.4283556979968262
    out = _laplacian(x.to(cuda), torch.tensor(h, dtype=torch.float64))
    ref = ref_torch.laplacian(x, torch.tensor(h, dtype=torch.float64))
    assert out.dtype == torch.float64
    assert rel(out.cpu().numpy(), ref.numpy()) < 1e-15


def test_laplacian_rejects(cuda):
    with pytest.raises(ValueError):
        _laplacian(torch.zeros(4, 5, device=cuda), 1.0)
    with pytest.raises(NotImplementedError):
        _laplacian(torch.zeros(1, 4, 5, device=cuda, dtype=torch.float16), 1.0)
