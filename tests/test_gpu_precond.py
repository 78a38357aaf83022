"""Learned preconditioner on the GPU: the network's forward pass and parameter gradients
against the CPU restatement (oracle/unet_oracle.py), then training on pairs that the
mixed-precision solver logs and the flexible GMRES that uses the trained network.

The CPU restatement rounds to bfloat16 where the GPU stores bfloat16. The remaining
difference is float32 MFMA accumulation against float64: a value that lands next to a
bfloat16 rounding boundary can round one way on the GPU and the other on the CPU. The bar
is therefore normwise: 1e-2 on the network output and on every parameter tensor's gradient.
"""
import dataclasses

import numpy as np
import pytest

from conftest import relerr
from oracle import unet_oracle as uo
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

TOL_NET = 1e-2


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = uo.init_params(seed)
    for name, o, sh in uo.param_layout()[0]:
        k = int(np.prod(sh))
        if name.endswith(".b"):
            p[o:o + k] = rng.normal(0, 0.05, k)
        if name == "head.W":
            p[o:o + k] = rng.normal(0, 0.3, k)
    return p.astype(np.float32)


@pytest.mark.parametrize("grid", [(8, 8, 8), (16, 16, 64), (32, 32, 32)])
def test_network_forward_and_gradients_match_cpu(grid):
    """8^3: every convolution on the voxel-row kernel (the coarse levels are narrower than a
    16-voxel row); 16 x 16 x 64: every convolution, forward and adjoint, on the LDS brick
    kernel (levels 16 x 16 x 64, 8 x 8 x 32, 4 x 4 x 16); 32^3 (the size of the small training
    box): the 16 -> 16 weight gradients on the brick kernel with more workgroups than the
    MFMA kernel would use, which the partial-sum buffer must hold."""
    torch = _torch()
    w = wl.mini(n=grid, ppc=4, seed=11)
    d = w.n
    V = int(np.prod(d))
    print("case: network on a %dx%dx%d grid, random parameters" % d, flush=True)
    sim = Simulation(w, precision="fp32", solver="mixed", stop_rule="true_residual")
    sim.precond_create()
    p = rand_params(3)
    mu_scale = 0.02
    sim.set_precond_params(p, mu_scale)
    rng = np.random.default_rng(5)
    r = rng.normal(size=(3, V)).astype(np.float32)
    mu = rng.normal(scale=0.02, size=(9, V)).astype(np.float32)
    e = (r + 0.2 * rng.normal(size=(3, V))).astype(np.float32)
    tr, tm, te = (torch.from_numpy(a).cuda() for a in (r, mu, e))
    tz = torch.zeros_like(tr)
    sim.precond_apply(tr.data_ptr(), tm.data_ptr(), tz.data_ptr())
    torch.cuda.synchronize()
    z_gpu = tz.cpu().numpy().astype(np.float64)
    net = uo.Net(p)
    z_cpu = net.forward(r.reshape((3,) + d), mu.reshape((9,) + d), mu_scale).reshape(3, V)
    err_z = relerr(z_gpu - r, z_cpu - r)
    loss_gpu, g_gpu = sim.precond_gradient(tr.data_ptr(), tm.data_ptr(), te.data_ptr())
    loss_cpu, g_cpu = uo.Net(p).loss_and_grad(r.reshape((3,) + d), mu.reshape((9,) + d), mu_scale, e.reshape((3,) + d))
    errs = {"z": err_z, "loss": abs(loss_gpu - loss_cpu) / loss_cpu}
    for name, o, sh in uo.param_layout()[0]:
        k = int(np.prod(sh))
        errs[name] = relerr(g_gpu[o:o + k], g_cpu[o:o + k])
    print("network gpu vs cpu:", {k: "%.1e" % v for k, v in errs.items()})
    for k, v in errs.items():
        assert v <= TOL_NET, (k, v)
    sim.close()


def _solve_pair(sim, torch, rhs_ptr, n):
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    return sim.gmres_solve(rhs_ptr, x.data_ptr())


def learned_case(torch, dt, log_steps=40, epochs=30):
    """Logs pairs from `log_steps` steps of a 16^3 box at time step dt, trains, then solves
    the next 5 steps' Maxwell systems with and without the network."""
    w = dataclasses.replace(wl.mini(n=(16, 16, 16), ppc=8, seed=9), dt=dt)
    st = wl.make_initial_state(w, perturb=0.05)
    print("case: 16^3 box, 8 ppc per species, dt = %g, learned preconditioner" % dt, flush=True)
    sim = Simulation(w, precision="fp32", solver="mixed", stop_rule="true_residual")
    sim.load_state(st)
    sim.initialize()
    sim.precond_create()
    sim.precond_log(2000)
    for _ in range(log_steps):
        sim.step()
    npairs = sim.precond_log(-1)
    _, identity_loss = sim.precond_train(epochs=0, batch=8, lr=2e-3, holdout=0.15)  # untrained net: z = r
    tl, hl = sim.precond_train(epochs=epochs, batch=8, lr=2e-3, holdout=0.15)
    sim.precond_log(0)
    plain, learned = [], []
    for _ in range(5):
        sim.updateElectroMagneticParameters()
        sim.evaluateMaxwellEquationMatrix()
        rhs_ptr, n = sim.device_array("rhs")
        sim.precond_use(False)
        plain.append(_solve_pair(sim, torch, rhs_ptr, n))
        sim.precond_use(True)
        learned.append(_solve_pair(sim, torch, rhs_ptr, n))
        sim.precond_use(False)
        sim.evaluateElectricField()
        sim.evaluateMagneticField()
        sim.moveParticles()
        sim.updateFields()
    sim.close()
    out = dict(pairs=npairs, identity_loss=identity_loss, train_loss=tl, holdout_loss=hl,
               plain_steps=np.mean([s["iterations"] for s in plain]), learned_steps=np.mean([s["iterations"] for s in learned]),
               plain_ms=1e3 * np.mean([s["seconds"] for s in plain]), learned_ms=1e3 * np.mean([s["seconds"] for s in learned]),
               fallbacks=sum(s["fallbacks"] for s in learned),
               worst_residual=max(s["true_residual"] for s in plain + learned))
    print("dt %g: pairs %d; held-out loss identity %.3e -> trained %.3e (train %.3e); held-out solves: plain %.1f Arnoldi "
          "steps %.3f ms, learned %.1f steps %.3f ms, fallbacks %d" % (
              dt, out["pairs"], identity_loss, hl, tl, out["plain_steps"], out["plain_ms"], out["learned_steps"],
              out["learned_ms"], out["fallbacks"]))
    return out


def test_trained_preconditioner_at_the_configured_time_step():
    """dt = 0.1 / omega_pe (the benchmark's step): A is close to the identity and plain GMRES
    needs about 4 Arnoldi steps, so there is little to gain; the network must still learn
    (held-out loss below the identity's) and the solves must reach the 1e-8 residual."""
    out = learned_case(_torch(), 0.1)
    assert out["worst_residual"] <= 1e-8
    assert out["holdout_loss"] < 0.5 * out["identity_loss"]


def test_trained_preconditioner_gains_on_a_stiffer_system():
    """dt = 0.6 / omega_pe (c dt / dx = 0.6): the curl-curl term matters, plain GMRES needs
    about 12 Arnoldi steps, and the trained network must cut them (measured: about 8)."""
    out = learned_case(_torch(), 0.6)
    assert out["worst_residual"] <= 1e-8
    assert out["holdout_loss"] < 0.5 * out["identity_loss"]
    assert out["learned_steps"] < out["plain_steps"]


def test_bad_network_falls_back_to_plain_gmres():
    """A network with huge random head weights is useless as a preconditioner; the solver
    must notice (fallbacks > 0 or it converges anyway) and still reach the 1e-8 residual."""
    torch = _torch()
    w = wl.mini(n=(8, 8, 8), ppc=4, seed=13)
    st = wl.make_initial_state(w, perturb=0.05)
    print("case: 8^3 box, deliberately poor network", flush=True)
    sim = Simulation(w, precision="fp32", solver="mixed", stop_rule="true_residual")
    sim.load_state(st)
    sim.initialize()
    sim.precond_create()
    p = rand_params(7)
    lay = {name: (o, sh) for name, o, sh in uo.param_layout()[0]}
    o, sh = lay["head.W"]
    p[o:o + int(np.prod(sh))] *= 1e3
    sim.set_precond_params(p, 0.01)
    sim.updateElectroMagneticParameters()
    sim.evaluateMaxwellEquationMatrix()
    rhs_ptr, n = sim.device_array("rhs")
    sim.precond_use(True)
    res = _solve_pair(sim, torch, rhs_ptr, n)
    print("poor network: %d Arnoldi steps, %d fallbacks, true residual %.2e" % (res["iterations"], res["fallbacks"],
                                                                             res["true_residual"]))
    assert res["true_residual"] <= 1e-8
    sim.close()
