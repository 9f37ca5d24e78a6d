"""Proposer network kernels (csrc/mlp.hip) against the numpy restatement
(oracle/mlp_ref.py, pinned by finite differences): forward pass, loss and
gradient, Adam step, in FP32 (relative tolerance 1e-4 on the gradient,
1e-5 on forward values), and a short training run that fits a smooth map."""
import numpy as np
import pytest

from fmm2d_amd import _native as nat
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(nat.torch_device())


def test_forward_and_gradient_match_oracle():
    import torch
    L = nat.lib()
    nin, nout, batch = 20, 17, 300
    rng = np.random.default_rng(3)
    p = M.init(nin, nout, seed=4) + 0.02 * rng.standard_normal(M.layout(nin, nout)[1])
    p = p.astype(np.float32).astype(np.float64)
    x = rng.standard_normal((batch, nin)).astype(np.float32).astype(np.float64)
    t = rng.standard_normal((batch, nout)).astype(np.float32).astype(np.float64)
    w = rng.uniform(0.1, 1.0, (batch, nout)).astype(np.float32).astype(np.float64)
    n = L.fmm2d_mlp_nparams(nin, nout)
    assert n == len(p)
    dp, dx, dt, dw = _dev(p), _dev(x), _dev(t), _dev(w)
    y = torch.zeros((batch, nout), dtype=torch.float32, device=dp.device)
    assert L.fmm2d_mlp_forward(nin, nout, nat.ptr(dp), batch, nat.ptr(dx), nat.ptr(y),
                               nat.stream_ptr()) == 0
    yref, _ = M.forward(p, x, nin, nout)
    assert np.allclose(y.cpu().numpy(), yref, rtol=1e-5, atol=1e-5)
    g = torch.zeros(n, dtype=torch.float32, device=dp.device)
    loss = torch.zeros(1, dtype=torch.float32, device=dp.device)
    assert L.fmm2d_mlp_loss_grad(nin, nout, nat.ptr(dp), batch, nat.ptr(dx), nat.ptr(dt),
                                 nat.ptr(dw), nat.ptr(g), nat.ptr(loss), nat.stream_ptr()) == 0
    lref, gref = M.loss_grad(p, x, t, w, nin, nout)
    assert abs(float(loss.cpu()) - lref) <= 1e-5 * abs(lref)
    gg = g.cpu().numpy()
    assert np.linalg.norm(gg - gref) <= 1e-4 * np.linalg.norm(gref)


def test_adam_matches_oracle():
    import torch
    L = nat.lib()
    rng = np.random.default_rng(5)
    n = 1000
    p = rng.standard_normal(n)
    m = 0.1 * rng.standard_normal(n)
    v = rng.uniform(0, 0.1, n)
    g = rng.standard_normal(n)
    dp, dm, dv, dg = _dev(p), _dev(m), _dev(v), _dev(g)
    assert L.fmm2d_mlp_adam(n, nat.ptr(dp), nat.ptr(dg), nat.ptr(dm), nat.ptr(dv), 1e-2, 0.9,
                            0.999, 1e-8, 7, nat.stream_ptr()) == 0
    pr, mr, vr = M.adam(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 7)
    assert np.allclose(dp.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)
    assert np.allclose(dm.cpu().numpy(), mr, rtol=1e-5, atol=1e-7)


def test_training_fits_smooth_map():
    """400 Adam steps on the device fit y = sin(x0) + x1 x2 to a small loss."""
    import torch
    L = nat.lib()
    nin, nout, batch = 3, 1, 512
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (batch, nin))
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2])[:, None]
    n = L.fmm2d_mlp_nparams(nin, nout)
    p = _dev(M.init(nin, nout, seed=1))
    m = torch.zeros(n, dtype=torch.float32, device=p.device)
    v = torch.zeros_like(m)
    g = torch.zeros_like(m)
    loss = torch.zeros(1, dtype=torch.float32, device=p.device)
    dx, dt = _dev(x), _dev(t)
    first = None
    for step in range(1, 401):
        assert L.fmm2d_mlp_loss_grad(nin, nout, nat.ptr(p), batch, nat.ptr(dx), nat.ptr(dt),
                                     None, nat.ptr(g), nat.ptr(loss), nat.stream_ptr()) == 0
        if first is None:
            first = float(loss.cpu())
        assert L.fmm2d_mlp_adam(n, nat.ptr(p), nat.ptr(g), nat.ptr(m), nat.ptr(v), 3e-3, 0.9,
                                0.999, 1e-8, step, nat.stream_ptr()) == 0
    last = float(loss.cpu())
    assert last < 0.02 * first and last < 5e-3, (first, last)
