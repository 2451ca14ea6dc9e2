"""Pins the network's CPU reference (oracle/mlp_ref.py) that the GPU training
and inference kernels are checked against.

- its BF16 rounding equals torch's float32 -> bfloat16 conversion (round to
  nearest even) bit for bit, ties, subnormals and specials included;
- with the roundings taken out, its loss gradient equals central finite
  differences of its own loss in float64 (a dropped term, a transposed
  operand or a wrong activation derivative fails this);
- the SiLU derivative equals finite differences of SiLU.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_ref


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 20000),
        # exact ties between two BF16 values (low 16 bits = 0x8000), both parities
        (np.arange(1, 2001, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 1e-40, -1e-40, 3.4e38], dtype=np.float32),
    ])
    got = mlp_ref.bf16(x).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    nan = mlp_ref.bf16(np.array([np.nan], dtype=np.float32))
    assert np.isnan(nan).all()


def test_silu_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (mlp_ref.silu(z + h) - mlp_ref.silu(z - h)) / (2 * h)
    assert np.allclose(mlp_ref.silu_grad(z), fd, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("dims", [(3, 5, 4), (4, 6, 6, 2)])
def test_gradient_without_rounding_matches_finite_differences(monkeypatch, dims):
    monkeypatch.setattr(mlp_ref, "bf16", lambda a: np.asarray(a, dtype=np.float64))
    rng = np.random.default_rng(len(dims))
    n_par = sum(i * o + o for i, o in zip(dims[:-1], dims[1:]))
    p = 0.5 * rng.standard_normal(n_par)
    rows = 7
    x = rng.standard_normal((rows, dims[0]))
    t = np.concatenate([np.zeros((rows, 1)), rng.standard_normal((rows, dims[-1]))], axis=1)
    im, isc = rng.standard_normal(dims[0]), 0.5 + rng.random(dims[0])
    om, osc = rng.standard_normal(dims[-1]), 0.5 + rng.random(dims[-1])
    # keep the targets' float32 rounding out of the comparison
    t[:, 1:] = (((t[:, 1:] - om) / osc).astype(np.float32).astype(np.float64)) * osc + om
    _, g = mlp_ref.loss_and_grads(p, dims, x, t, im, isc, om, osc)
    h = 1e-6
    fd = np.zeros_like(p)
    for k in range(n_par):
        e = np.zeros_like(p)
        e[k] = h
        lp, _ = mlp_ref.loss_and_grads(p + e, dims, x, t, im, isc, om, osc)
        lm, _ = mlp_ref.loss_and_grads(p - e, dims, x, t, im, isc, om, osc)
        fd[k] = (lp - lm) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), np.abs(g - fd).max()
