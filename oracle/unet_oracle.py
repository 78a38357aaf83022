"""CPU restatement (numpy) of the learned Maxwell preconditioner: a 3-D convolution U-Net.

TEST INFRASTRUCTURE ONLY: the checker of the HIP network kernels (forward, loss and
backward). Only tests/ may import it; the product path never does.

The reference has no preconditioner (its GMRES is unpreconditioned,
PIC++/specialmath.cpp:367-671), so there is no reference file to follow. This module
defines the network exactly as the HIP kernels compute it, including where they round to
bfloat16, so that GPU gradients can be compared element by element. With `exact=True`
the rounding is switched off and everything runs in float64. That mode is what the
finite-difference checks pin.

Network (grid dims divisible by 4, periodic in every axis, channels last):
  x0 = pack(r / s, mu / mu_scale)           16 ch: 3 residual, 9 dielectric tensor, 4 zero
  e1 = relu(conv3(x0))      16 -> 16        level 0
  e2 = relu(conv3(pool(e1)))  16 -> 32      level 1
  e3 = relu(conv3(pool(e2)))  32 -> 64      level 2
  d3 = relu(conv3(e3))      64 -> 32        level 2
  d2 = relu(conv3(up(d3) + e2)) 32 -> 16    level 1
  d1 = relu(conv3(up(d2) + e1)) 16 -> 16    level 0
  y  = d1 Wh^T + bh          16 -> 3 (1x1x1, float32 weights)
  z  = s (r / s + y) = r + s y,   s = ||r|| / sqrt(3 V)
Loss per sample: ||z - e||^2 / ||e||^2 for the target correction e.
conv3: y[v, co] = sum_t sum_ci x[v + o_t, ci] W[t, co, ci] + b[co], with taps
t = 9 (dx+1) + 3 (dy+1) + (dz+1), o_t = (dx, dy, dz) in {-1, 0, 1}^3.
pool: mean over 2x2x2 children; up: copy to the 8 children.
"""
from __future__ import annotations

import numpy as np

CONVS = [("enc1", 16, 16), ("enc2", 16, 32), ("enc3", 32, 64), ("dec3", 64, 32), ("dec2", 32, 16), ("dec1", 16, 16)]
HEAD = (16, 3)
TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def param_layout():
    """(name, offset, shape) of every parameter in the flat float32 vector shared with the GPU."""
    out, off = [], 0
    for name, ci, co in CONVS:
        out.append((name + ".W", off, (27, co, ci)))
        off += 27 * co * ci
        out.append((name + ".b", off, (co,)))
        off += co
    out.append(("head.W", off, (HEAD[1], HEAD[0])))
    off += HEAD[0] * HEAD[1]
    out.append(("head.b", off, (HEAD[1],)))
    off += HEAD[1]
    return out, off


def unpack_params(flat):
    lay, n = param_layout()
    assert flat.size == n
    return {name: flat[o:o + int(np.prod(sh))].reshape(sh) for name, o, sh in lay}


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and return them as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).reshape(np.shape(x))


class Net:
    def __init__(self, flat, exact=False):
        self.p = unpack_params(np.asarray(flat, np.float64 if exact else np.float32))
        self.exact = exact

    def q(self, x):
        """The storage rounding of an activation or a bf16 operand (none when exact)."""
        return np.asarray(x, np.float64) if self.exact else bf16(np.asarray(x, np.float32)).astype(np.float64)

    # ------------------------------------------------------------ building blocks
    def conv3(self, x, W, b=None):
        Wq = self.q(W)
        y = np.zeros(x.shape[:3] + (W.shape[1],))
        for t, (dx, dy, dz) in enumerate(TAPS):
            xs = np.roll(x, shift=(-dx, -dy, -dz), axis=(0, 1, 2))  # xs[v] = x[v + o_t]
            y += xs @ Wq[t].T
        if b is not None:
            y += np.asarray(b, np.float64)
        return y

    def conv3_adjoint(self, dz, W):
        """dx[u] = sum_t dz[u - o_t] W[t] (the transpose of conv3 in its input)."""
        Wq = self.q(W)
        dx = np.zeros(dz.shape[:3] + (W.shape[2],))
        for t, (dx_, dy_, dz_) in enumerate(TAPS):
            ds = np.roll(dz, shift=(dx_, dy_, dz_), axis=(0, 1, 2))  # ds[u] = dz[u - o_t]
            dx += ds @ Wq[t]
        return dx

    def conv3_wgrad(self, dz, x):
        """dW[t] = sum_v dz[v] (outer) x[v + o_t]; db = sum_v dz[v]."""
        co, ci = dz.shape[3], x.shape[3]
        dW = np.zeros((27, co, ci))
        dzf = dz.reshape(-1, co)
        for t, (dx, dy, dz_) in enumerate(TAPS):
            xs = np.roll(x, shift=(-dx, -dy, -dz_), axis=(0, 1, 2)).reshape(-1, ci)
            dW[t] = dzf.T @ xs
        return dW, dzf.sum(axis=0)

    @staticmethod
    def pool(x):
        d0, d1, d2, c = x.shape
        return x.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).mean(axis=(1, 3, 5))

    @staticmethod
    def pool_sum(x):
        d0, d1, d2, c = x.shape
        return x.reshape(d0 // 2, 2, d1 // 2, 2, d2 // 2, 2, c).sum(axis=(1, 3, 5))

    @staticmethod
    def up(x):
        return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)

    # ------------------------------------------------------------ forward / backward
    def forward(self, r, mu, mu_scale):
        """r: (3, d0, d1, d2) residual planes, mu: (9, d0, d1, d2). Returns z (3, d0, d1, d2)."""
        p, q = self.p, self.q
        r = np.asarray(r, np.float64)
        V = r[0].size
        s = np.sqrt(np.sum(r * r) / (3 * V))
        self.s = s
        x0 = np.zeros(r.shape[1:] + (16,))
        if s > 0:
            x0[..., 0:3] = np.moveaxis(r / s, 0, -1)
        x0[..., 3:12] = np.moveaxis(np.asarray(mu, np.float64) / mu_scale, 0, -1)
        a = {"x0": q(x0)}
        relu = lambda v: np.maximum(v, 0.0)
        a["e1"] = q(relu(self.conv3(a["x0"], p["enc1.W"], p["enc1.b"])))
        a["p1"] = q(self.pool(a["e1"]))
        a["e2"] = q(relu(self.conv3(a["p1"], p["enc2.W"], p["enc2.b"])))
        a["p2"] = q(self.pool(a["e2"]))
        a["e3"] = q(relu(self.conv3(a["p2"], p["enc3.W"], p["enc3.b"])))
        a["d3"] = q(relu(self.conv3(a["e3"], p["dec3.W"], p["dec3.b"])))
        a["u2"] = q(self.up(a["d3"]) + a["e2"])
        a["d2"] = q(relu(self.conv3(a["u2"], p["dec2.W"], p["dec2.b"])))
        a["u1"] = q(self.up(a["d2"]) + a["e1"])
        a["d1"] = q(relu(self.conv3(a["u1"], p["dec1.W"], p["dec1.b"])))
        Wh = np.asarray(p["head.W"], np.float64)
        y = a["d1"] @ Wh.T + np.asarray(p["head.b"], np.float64)  # (d0, d1, d2, 3)
        self.a = a
        self.y = y
        self.r = r
        return r + s * np.moveaxis(y, -1, 0)

    def loss_and_grad(self, r, mu, mu_scale, e):
        """Loss ||z - e||^2 / ||e||^2 and its gradient as a flat vector (param_layout order)."""
        z = self.forward(r, mu, mu_scale)
        e = np.asarray(e, np.float64)
        ee = np.sum(e * e)
        diff = z - e
        loss = np.sum(diff * diff) / ee
        dy = np.moveaxis(self.s * 2.0 * diff / ee, 0, -1)  # (d0, d1, d2, 3)
        p, a, q = self.p, self.a, self.q
        g = {}
        g["head.W"] = dy.reshape(-1, 3).T @ a["d1"].reshape(-1, 16)
        g["head.b"] = dy.reshape(-1, 3).sum(axis=0)
        dd1 = dy @ np.asarray(p["head.W"], np.float64)

        def layer(name, dout, act, inp, need_input_grad=True):
            dz = q(dout * (act > 0))
            g[name + ".W"], g[name + ".b"] = self.conv3_wgrad(dz, inp)
            return self.conv3_adjoint(dz, p[name + ".W"]) if need_input_grad else None

        du1 = layer("dec1", dd1, a["d1"], a["u1"])
        du2 = layer("dec2", self.pool_sum(du1), a["d2"], a["u2"])
        de3 = layer("dec3", self.pool_sum(du2), a["d3"], a["e3"])
        dp2 = layer("enc3", de3, a["e3"], a["p2"])
        de2 = self.up(dp2) / 8.0 + du2
        dp1 = layer("enc2", de2, a["e2"], a["p1"])
        de1 = self.up(dp1) / 8.0 + du1
        layer("enc1", de1, a["e1"], a["x0"], need_input_grad=False)
        lay, n = param_layout()
        flat = np.zeros(n)
        for name, o, sh in lay:
            flat[o:o + int(np.prod(sh))] = np.asarray(g[name]).reshape(-1)
        return loss, flat


def init_params(seed=0):
    """He-normal initial weights (float32), zero biases; the head starts at zero, so the
    untrained preconditioner is the identity (z = r)."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout()
    flat = np.zeros(n, np.float32)
    for name, o, sh in lay:
        if name.endswith(".W") and not name.startswith("head"):
            fan_in = sh[0] * sh[2]
            flat[o:o + int(np.prod(sh))] = rng.normal(0.0, np.sqrt(2.0 / fan_in), int(np.prod(sh)))
    return flat
