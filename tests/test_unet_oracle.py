"""Pins the CPU restatement of the learned preconditioner network (oracle/unet_oracle.py),
which the GPU gradient checks compare against. The network has no reference counterpart,
so each check below is an independent one:
- the 3x3x3 periodic convolution against scipy's `correlate` in wrap mode;
- the convolution adjoint and the pool/upsample pair by the inner-product identity;
- the whole backward pass against central finite differences (float64, no rounding);
- the bfloat16 rounding against torch's float32 -> bfloat16 conversion.
A transposed weight, a flipped tap, a wrong skip connection or a missing 1/8 in the pool
gradient fails one of them.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import unet_oracle as uo


def rand_params(seed, scale_head=0.3):
    rng = np.random.default_rng(seed)
    p = uo.init_params(seed).astype(np.float64)
    lay, n = uo.param_layout()
    for name, o, sh in lay:
        k = int(np.prod(sh))
        if name.endswith(".b"):
            p[o:o + k] = rng.normal(0, 0.05, k)
        if name == "head.W":
            p[o:o + k] = rng.normal(0, scale_head, k)
    return p


def test_conv3_matches_scipy_correlate_wrap():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(6, 5, 4, 3))
    W = rng.normal(size=(27, 2, 3))
    net = uo.Net(np.zeros(uo.param_layout()[1]), exact=True)
    y = net.conv3(x, W)
    ref = np.zeros((6, 5, 4, 2))
    for co in range(2):
        for ci in range(3):
            k = W[:, co, ci].reshape(3, 3, 3)  # t = 9 (dx+1) + 3 (dy+1) + (dz+1)
            ref[..., co] += ndi.correlate(x[..., ci], k, mode="wrap")
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv3_adjoint_and_pool_up_identities():
    rng = np.random.default_rng(2)
    net = uo.Net(np.zeros(uo.param_layout()[1]), exact=True)
    x = rng.normal(size=(8, 4, 6, 5))
    W = rng.normal(size=(27, 7, 5))
    yv = rng.normal(size=(8, 4, 6, 7))
    lhs = np.sum(net.conv3(x, W) * yv)
    rhs = np.sum(x * net.conv3_adjoint(yv, W))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    c = rng.normal(size=(4, 2, 3, 5))
    assert abs(np.sum(net.pool(x) * c) - np.sum(x * net.up(c)) / 8.0) <= 1e-10 * np.sum(np.abs(x))
    assert np.allclose(net.pool_sum(net.up(c)), 8 * c)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4000).astype(np.float32) * 10 ** rng.uniform(-6, 6, 4000).astype(np.float32),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(uo.bf16(x), ref)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    d = (8, 8, 4)
    p = rand_params(5)
    r = rng.normal(size=(3,) + d)
    mu = rng.normal(scale=0.01, size=(9,) + d)
    e = r + 0.1 * rng.normal(size=(3,) + d)
    net = uo.Net(p, exact=True)
    loss, g = net.loss_and_grad(r, mu, 0.01, e)
    lay, n = uo.param_layout()
    picks = []
    for name, o, sh in lay:  # two entries of every parameter tensor
        k = int(np.prod(sh))
        picks += [o + int(i) for i in rng.choice(k, size=2, replace=False)]
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (uo.Net(pp, exact=True).loss_and_grad(r, mu, 0.01, e)[0] - uo.Net(pm, exact=True).loss_and_grad(r, mu, 0.01, e)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3), (i, fd, g[i])


def test_untrained_network_is_identity():
    rng = np.random.default_rng(6)
    d = (4, 4, 4)
    r = rng.normal(size=(3,) + d)
    z = uo.Net(uo.init_params(0)).forward(r, rng.normal(size=(9,) + d), 1.0)
    assert np.allclose(z, r)
