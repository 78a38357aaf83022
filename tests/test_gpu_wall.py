"""Reflecting-wall configuration (the shock setup's boundaries) on the FP64 HIP path, phase
by phase, against golden vectors of the compiled reference (tests/golden/wall_ref.npz, made
by tests/golden/make_golden.py from PIC++ built from its own sources with the
SUPER_CONDUCTOR_LEFT boundary and a ConstantBoundaryFieldEvaluator at the right end).

The wall case: 12x5x5 cells, a gamma 2 e-/p+ flow along -x onto the wall at x = L, dt = 0.5,
5 steps, perturbed fields, no injection. The bar is the FP64 path's: normwise relative error
<= 1e-12 on particle state, J, rho, mu, every field and the energies.
"""
import os

import numpy as np
import pytest

from conftest import GoldenState, relerr
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))


def wall_workload():
    from picpp_amd import workloads as wl

    return wl.shock(n=(12, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=5, seed=19, name="wall")


def compare_particles(sim, ref):
    g = sim.particles()
    gi = np.argsort(g["ids"])
    r = ref[np.argsort(ref[:, 0])]
    assert np.array_equal(g["ids"][gi].astype(np.float64), r[:, 0])
    ex = relerr(g["x"][gi], r[:, 1:4])
    ep = relerr(g["p"][gi], r[:, 4:7])
    return ex, ep


def test_wall_steps_match_reference():
    d = np.load(os.path.join(HERE, "golden", "wall_ref.npz"))
    w = wall_workload()
    st = GoldenState(d)
    print("case", w.name, "particles", len(st.species), flush=True)
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
    print("wall vs reference (max over steps):", {k: "%.1e" % v for k, v in errs.items()})
    sim.close()
