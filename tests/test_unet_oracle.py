"""Pins the FP64 U-Net restatement (oracle/unet_ref.py) that checks the GPU
network: convolution against torch's conv2d, every parameter gradient against
central finite differences, and the bf16 rounding against torch's conversion."""
import numpy as np
import pytest
import torch

from oracle import unet_ref as U


def test_conv3x3_matches_torch_conv2d():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 9, 13, 5))
    w = rng.standard_normal((7, 3, 3, 5))
    b = rng.standard_normal(7)
    ref = torch.nn.functional.conv2d(torch.from_numpy(x).permute(0, 3, 1, 2), torch.from_numpy(w).permute(0, 3, 1, 2),
                                     torch.from_numpy(b), padding=1).permute(0, 2, 3, 1).numpy()
    np.testing.assert_allclose(U.conv3x3(x, w, b), ref, rtol=1e-12, atol=1e-12)


def test_conv3x3_gradients_match_autograd():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 10, 4))
    w = rng.standard_normal((3, 3, 3, 4))
    g = rng.standard_normal((2, 6, 10, 3))
    xt = torch.from_numpy(x).permute(0, 3, 1, 2).requires_grad_()
    wt = torch.from_numpy(w).permute(0, 3, 1, 2).detach().requires_grad_()
    bt = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.conv2d(xt, wt, bt, padding=1)
    y.backward(torch.from_numpy(g).permute(0, 3, 1, 2))
    dw, db = U.conv3x3_wgrad(g, x)
    np.testing.assert_allclose(dw, wt.grad.permute(0, 2, 3, 1).numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(U.conv3x3_dgrad(g, w), xt.grad.permute(0, 2, 3, 1).numpy(), rtol=1e-10, atol=1e-10)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e-39]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).double().numpy()
    np.testing.assert_array_equal(U.bf16_round(v), ref)


def test_layout_counts():
    assert U.param_count(32, 4) == sum(9 * ci * co + co for _, ci, co, _, _ in U.layout(32, 4))
    kinds = [k for k, *_ in U.layout(32, 4)]
    assert kinds[0] == U.EDGE_IN and kinds[-1] == U.EDGE_OUT and kinds.count(U.CONV) == 13 and len(kinds) == 15


@pytest.mark.parametrize("base,levels", [(2, 3), (1, 2)])
def test_gradients_match_finite_differences(base, levels):
    net = U.UnetRef(base, levels)
    rng = np.random.default_rng(10 + levels)
    H = W = 4 << (levels - 1)
    x = rng.standard_normal((2, H, W))
    t = x + 0.3 * rng.standard_normal((2, H, W))
    p = U.init_params(base, levels, 5)
    loss, g = net.loss_grad(p, x, t, 0.7, rnd=None)
    idx = rng.choice(net.n, size=min(net.n, 400), replace=False)
    eps = 1e-6
    fd = np.empty(len(idx))
    for k, i in enumerate(idx):
        q = p.copy()
        q[i] += eps
        lp = net.loss_grad(q, x, t, 0.7, rnd=None)[0]
        q[i] -= 2 * eps
        lm = net.loss_grad(q, x, t, 0.7, rnd=None)[0]
        fd[k] = (lp - lm) / (2 * eps)
    err = np.abs(fd - g[idx])
    assert np.max(err) <= 1e-6 * np.max(np.abs(g)) + 1e-5 * np.max(np.abs(fd)), np.max(err)
    # a dropped layer gradient would show as zeros where differences are not
    nz = np.abs(fd) > 1e-8
    assert nz.sum() > 0.5 * len(idx)


def test_bf16_emulation_is_close_to_exact():
    net = U.UnetRef(32, 3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 16, 16))
    p = U.init_params(32, 3, 1)
    y0, _ = net.forward(p, x, 1.0, rnd=None)
    y1, _ = net.forward(p, x, 1.0)
    rel = np.linalg.norm(y1 - y0) / np.linalg.norm(y0 - x)
    assert 1e-5 < rel < 5e-2
