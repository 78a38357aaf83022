"""Reflecting-wall configuration (the shock setup's boundaries) on the FP64 HIP path, phase
by phase, against golden vectors of the compiled reference (tests/golden/wall_ref.npz, made
by tests/golden/make_golden.py from PIC++ built from its own sources with the
SUPER_CONDUCTOR_LEFT boundary and a ConstantBoundaryFieldEvaluator at the right end).

The wall case: 12x5x5 cells, a gamma 2 e-/p+ flow along -x onto the wall at x = L, dt = 0.5,
5 steps, perturbed fields, no injection. The escape case: the same with the right half of the
particles turned to move along +x; about 10% of them leave through the free right end over
the 5 steps and are removed. The bar is the FP64 path's: normwise relative error <= 1e-12 on
particle state, J, rho, mu, every field and the energies, and the same particles kept.

The FP32 form runs the same cases against the FP64 path (itself pinned above): the same
particles must survive and every quantity must agree to 2e-5 (the FP32 bar) over 5 steps of
dt = 0.5 (measured: about 1e-7).
"""
import os

import numpy as np
import pytest

from conftest import GoldenState, relerr
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL32 = 2e-5
HERE = os.path.dirname(os.path.abspath(__file__))


def wall_workload():
    from picpp_amd import workloads as wl

    return wl.shock(n=(12, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=5, seed=19, name="wall",
                    inject=False)


def compare_particles(sim, ref):
    g = sim.particles()
    gi = np.argsort(g["ids"])
    r = ref[np.argsort(ref[:, 0])]
    assert np.array_equal(g["ids"][gi].astype(np.float64), r[:, 0])
    ex = relerr(g["x"][gi], r[:, 1:4])
    ep = relerr(g["p"][gi], r[:, 4:7])
    return ex, ep


@pytest.mark.parametrize("fixture", ["wall_ref.npz", "wall_out_ref.npz"])
def test_wall_steps_match_reference(fixture):
    d = np.load(os.path.join(HERE, "golden", fixture))
    w = wall_workload()
    st = GoldenState(d)
    print("case", w.name, fixture, "particles", len(st.species), flush=True)
    sim = Simulation(w)
    sim.load_state(st)
    sim.initialize()
    errs = {}

    def check(name, got, want):
        e = relerr(got, want)
        errs[name] = max(errs.get(name, 0.0), e)
        assert e <= TOL, (name, e)

    check("energy", sim.updateEnergy(), d["s0/energy"])
    check("E", sim.array("E"), d["s0/E"])
    check("B", sim.array("B"), d["s0/B"])
    n = w.n
    for step in range(1, w.steps + 1):
        full = step <= 2
        sim.updateElectroMagneticParameters()
        if full:
            for name in ("J", "mu", "rhohat", "Pi", "divPi"):
                check(name, sim.array(name), d["s%d/%s" % (step, name)])
            sim.evaluateMaxwellEquationMatrix()
            rhs = d["s%d/rhs" % step][2:n[0] + 4, 2:n[1] + 2, 2:n[2] + 2]
            check("rhs", sim.array("rhs"), rhs)
        sim.evaluateElectricField()
        sim.evaluateMagneticField()
        if full:
            for name in ("Etheta", "Enew", "Bnew"):
                check(name, sim.array(name), d["s%d/%s" % (step, name)])
        sim.moveParticles()
        ex, ep = compare_particles(sim, d["s%d/particles" % step])
        errs["x"] = max(errs.get("x", 0.0), ex)
        errs["p"] = max(errs.get("p", 0.0), ep)
        assert ex <= TOL and ep <= TOL, (step, ex, ep)
        dens = sim.array("density")
        check("rho", dens[..., 5 * len(w.species)], d["s%d/rho" % step])
        if full:
            for s in range(len(w.species)):
                check("conc", dens[..., 5 * s], d["s%d/conc%d" % (step, s)])
                check("energy_density", dens[..., 5 * s + 1], d["s%d/energy%d" % (step, s)])
        sim.updateFields()
        check("E", sim.array("E"), d["s%d/E" % step])
        check("B", sim.array("B"), d["s%d/B" % step])
        check("energy", sim.updateEnergy(), d["s%d/energy" % step])
    print("wall vs reference (max over steps):", {k: "%.1e" % v for k, v in errs.items()},
          "particles left:", len(sim.particles()["ids"]))
    sim.close()


@pytest.mark.parametrize("fixture", ["wall_ref.npz", "wall_out_ref.npz"])
def test_fp32_wall_tracks_fp64(fixture):
    d = np.load(os.path.join(HERE, "golden", fixture))
    w = wall_workload()
    st = GoldenState(d)
    print("case fp32", w.name, fixture, flush=True)
    sims = {}
    for prec in ("fp64", "fp32"):
        sims[prec] = Simulation(w, precision=prec)
        sims[prec].load_state(st)
        sims[prec].initialize()
    errs = {}
    for step in range(1, w.steps + 1):
        for sim in sims.values():
            sim.step()
        a, b = sims["fp32"].particles(), sims["fp64"].particles()
        ai, bi = np.argsort(a["ids"]), np.argsort(b["ids"])
        assert np.array_equal(a["ids"][ai], b["ids"][bi]), step
        cur = {"x": relerr(a["x"][ai], b["x"][bi]), "p": relerr(a["p"][ai], b["p"][bi])}
        for name in ("E", "B", "J", "density"):
            cur[name] = relerr(sims["fp32"].array(name), sims["fp64"].array(name))
        cur["energy"] = relerr(sims["fp32"].updateEnergy()[:4], sims["fp64"].updateEnergy()[:4])
        for k, v in cur.items():
            errs[k] = max(errs.get(k, 0.0), v)
    print("fp32 vs fp64 on the wall (max over steps):", {k: "%.1e" % v for k, v in errs.items()},
          "particles left:", len(a["ids"]))
    for k, v in errs.items():
        assert v <= TOL32, (k, v)
    for sim in sims.values():
        sim.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_long_shock_run_tracks_reference(precision, tol):
    """150 steps of the shock set-up (gamma 10 flow onto the wall, 16x8x8 cells, no injection)
    against the compiled reference's energies and E, B, rho norms at every step and a particle
    sample at the end (tests/golden/wall_long_ref.npz). FP64 keeps to 1e-10 (measured about
    1e-12 after 150 steps); the FP32 form to 1e-5."""
    from picpp_amd import workloads as wl

    d = np.load(os.path.join(HERE, "golden", "wall_long_ref.npz"))
    w = wl.shock(n=(16, 8, 8), ppc=8, steps=150, inject=False, seed=5, name="wall-long")
    print("case", w.name, precision, flush=True)
    sim = Simulation(w, precision=precision)
    sim.load_state(GoldenState(d))
    sim.initialize()
    worst = {}

    def track(name, got, want):
        e = relerr(got, want)
        worst[name] = max(worst.get(name, 0.0), e)

    for step in range(1, w.steps + 1):
        sim.step()
        track("energy", sim.updateEnergy()[:4], d["s%d/energy" % step][:4])
        for f in ("E", "B"):
            a = sim.array(f)
            track(f, [np.sqrt(np.sum(a * a)), np.sum(a)], d["s%d/%s_norm" % (step, f)])
        rho = sim.array("density")[..., 5 * len(w.species)]
        track("rho", np.sqrt(np.sum(rho * rho)), d["s%d/rho_norm" % step][0])
    g = sim.particles()
    ref = d["final_particles_sample"]
    idx = {int(i): k for k, i in enumerate(g["ids"])}
    sel = np.array([idx[int(i)] for i in ref[:, 0]])
    track("x", g["x"][sel], ref[:, 1:4])
    track("p", g["p"][sel], ref[:, 4:7])
    print("%s over 150 steps vs reference (worst):" % precision, {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= tol, (k, v)
    sim.close()
