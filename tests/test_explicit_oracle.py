"""Pins the CPU restatement of the explicit (BUNEMAN / Tristan) step
(oracle/explicit_oracle.cpp), which the GPU explicit path is checked against:
- against the compiled reference's own explicit step (tests/golden/explicit_ref.npz, made by
  oracle/ref_harness_explicit.cpp): particles, Yee E and B after every step and the Yee
  current of the first two steps, to 1e-12 normwise;
- against what the mathematics fixes, independently of the reference: the zig-zag current
  obeys the plane-summed continuity equation for the linear node density exactly and sums to
  sum q w v / V, the B half steps keep div B at its initial value, and with E = 0 and uniform
  B the mover follows the closed form of its rotation.
"""
import os

import numpy as np

from conftest import relerr
from oracle.oracle import OracleSim
from picpp_amd import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    return np.load(os.path.join(HERE, "golden", "explicit_ref.npz"))


def explicit_workload():
    import dataclasses
    return dataclasses.replace(wl.mini(), dt=0.4, steps=6, name="explicit")


def load_oracle(d, w):
    class St:
        pass
    st = St()
    st.species, st.ids, st.x, st.p = d["in/species"], d["in/ids"], d["in/x"], d["in/p"]
    N = tuple(v + 4 for v in w.n)
    st.E = np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3))
    st.B = np.zeros(N + (3,))
    o = OracleSim(w, st)
    o.yee_init()
    for name in wl.YEE_NAMES:
        o.set_yee(name, d["in/" + name])
    o.yee_exchange()
    return o


def by_id(parts):
    return parts[np.argsort(parts[:, 0])]


def test_explicit_oracle_matches_reference():
    d = golden()
    w = explicit_workload()
    o = load_oracle(d, w)
    worst = {}
    for name in wl.YEE_NAMES:
        worst["%s0" % name] = relerr(o.yee(name), d["s0/" + name])
    for step in range(1, w.steps + 1):
        o.explicit_step()
        p = o.particles()
        mine = by_id(np.column_stack([p["ids"].astype(float), p["x"], p["p"]]))
        ref = by_id(d["s%d/particles" % step])
        assert np.array_equal(mine[:, 0], ref[:, 0]), step
        for key, sl in (("x", slice(1, 4)), ("p", slice(4, 7))):
            worst[key] = max(worst.get(key, 0.0), relerr(mine[:, sl], ref[:, sl]))
        for name in wl.YEE_NAMES:
            worst[name] = max(worst.get(name, 0.0), relerr(o.yee(name), d["s%d/%s" % (step, name)]))
        if step <= 2:
            for name in ("Jx", "Jy", "Jz"):
                worst[name] = max(worst.get(name, 0.0), relerr(o.yee(name), d["s%d/%s" % (step, name)]))
    print("explicit oracle vs reference:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)


def line_density(w, ax, x, q_w):
    """Linear (tent) node charge density along axis ax, summed over the other two axes: the
    ghosted N + 1 nodes, folded periodically (node pairs (N - i, 4 - i), i = 0..4, as the
    reference folds node arrays), per unit length."""
    N = w.n[ax] + 4
    x0 = w.n[ax] * w.dx - 2 * w.dx
    rho = np.zeros(N + 1)
    r = (x[:, ax] - x0) / w.dx
    i0 = np.floor(r).astype(int)
    f = r - i0
    np.add.at(rho, i0, q_w * (1 - f))
    np.add.at(rho, i0 + 1, q_w * f)
    for i in range(5):
        rho[N - i] += rho[4 - i]
        rho[4 - i] = rho[N - i]
    return rho / w.dx ** 3


def test_zigzag_conserves_charge():
    """Summed over the two transverse axes, the zig-zag current obeys the 1-D continuity
    equation with the linear node density exactly (the transverse fluxes telescope out and
    the transverse weights sum to 1); summed over everything, each component is sum q w v / V.
    (The full 3-D continuity holds only up to the bilinear d_y d_z / 12 terms the midpoint
    weights leave out, which the reference's scheme does not carry.)"""
    d = golden()
    w = explicit_workload()
    o = load_oracle(d, w)
    m, q, wt, _ = wl.species_arrays(w)
    for step in range(3):
        o.tristan_b_half()
        o.tristan_move()  # particles moved, escapes listed but still in the vector
        p = o.particles()
        sp = p["species"]
        g = np.sqrt((p["p"] ** 2).sum(1) / m[sp] ** 2 + 1.0)
        v = p["p"] / (m[sp] * g)[:, None]
        prev = p["x"] - v * w.dt
        o.tristan_flux()
        qw = q[sp] * wt[sp]
        for ax, name in enumerate(("Jx", "Jy", "Jz")):
            J = o.yee(name)
            other = tuple(a for a in range(3) if a != ax)
            own_o = [slice(2, w.n[a] + 2) for a in other]
            Jl = np.moveaxis(J, ax, 0)[:, own_o[0], own_o[1]].sum(axis=(1, 2))  # faces along ax
            drho = (line_density(w, ax, p["x"], qw) - line_density(w, ax, prev, qw)) / w.dt
            res = np.abs(drho[2:w.n[ax] + 2] + (Jl[2:w.n[ax] + 2] - Jl[1:w.n[ax] + 1]) / w.dx).max()
            scale = np.abs(drho[2:w.n[ax] + 2]).max()
            total = Jl[2:w.n[ax] + 2].sum()
            want = (qw * v[:, ax]).sum() / w.dx ** 3
            print("step %d %s: 1-D continuity residual %.2e (of %.2e); total %.6e vs sum q w v / V %.6e"
                  % (step, name, res, scale, total, want))
            assert res <= 1e-12 * scale
            assert abs(total - want) <= 1e-12 * np.abs(qw * v[:, ax]).sum() / w.dx ** 3
        o.tristan_b_half()
        o.tristan_e()
        o.erase_and_exchange()
        o.yee_exchange()


def div_b(o, w):
    # cell (i, j, k): Bx on x faces i, i + 1 (N + 1 along x), By on y faces, Bz on z faces
    Bx, By, Bz = o.yee("Bx"), o.yee("By"), o.yee("Bz")
    return ((Bx[1:, :, :] - Bx[:-1, :, :]) + (By[:, 1:, :] - By[:, :-1, :]) + (Bz[:, :, 1:] - Bz[:, :, :-1])) / w.dx


def test_b_half_steps_keep_div_b():
    d = golden()
    w = explicit_workload()
    o = load_oracle(d, w)
    own = (slice(2, w.n[0] + 2),) * 1 + (slice(2, w.n[1] + 2), slice(2, w.n[2] + 2))
    d0 = div_b(o, w)[own]
    for _ in range(4):
        o.explicit_step()
    d1 = div_b(o, w)[own]
    change = np.abs(d1 - d0).max() / np.abs(o.yee("Bx")).max()
    print("div B change over 4 steps (relative to max |Bx|): %.2e" % change)
    assert change < 1e-13


def test_mover_rotation_closed_form():
    """With E = 0 and a uniform B the mover's update is a rotation about B followed by the
    reference's scaling: with t = beta |B| / gamma and f = 2 / sqrt(1 + t^2)
    (PIC++/mover.cpp:732-741; Boris's f = 2 / (1 + t^2) would keep |p| exactly), p along B is
    unchanged and |p_perp|^2 is multiplied by (1 - f t^2)^2 + f^2 t^2. A transposed cross
    product or a wrong beta changes p_par or this factor."""
    w = explicit_workload()
    d = golden()
    o = load_oracle(d, w)
    Bv = np.array([0.3, -0.7, 0.2])
    for name in ("Ex", "Ey", "Ez"):
        o.set_yee(name, np.zeros(o.yee_dims(name)))
    for name, b in zip(("Bx", "By", "Bz"), Bv):
        o.set_yee(name, np.full(o.yee_dims(name), b))
    before = o.particles()
    o.tristan_move()
    after = o.particles()
    m, q, _, _ = wl.species_arrays(w)
    sp = before["species"]
    e = Bv / np.linalg.norm(Bv)
    p0, p1 = before["p"], after["p"]
    g = np.sqrt((p0 ** 2).sum(1) / m[sp] ** 2 + 1.0)
    t = q[sp] * w.dt / (2.0 * m[sp]) / g * np.linalg.norm(Bv)
    f = 2.0 / np.sqrt(1.0 + t * t)
    par0, par1 = p0 @ e, p1 @ e
    perp0 = ((p0 - np.outer(par0, e)) ** 2).sum(1)
    perp1 = ((p1 - np.outer(par1, e)) ** 2).sum(1)
    factor = (1 - f * t * t) ** 2 + (f * t) ** 2
    err_par = np.abs(par1 - par0).max() / np.abs(p0).max()
    err_perp = np.abs(perp1 / perp0 - factor).max()
    print("rotation: p_par change %.2e, |p_perp|^2 factor error %.2e (factor - 1 up to %.2e)"
          % (err_par, err_perp, np.abs(factor - 1).max()))
    assert err_par < 1e-14 and err_perp < 1e-13
