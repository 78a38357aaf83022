"""Divergence cleanup of E and B (pic_cleanup_divergence, cleanupDivergence;
pic_cleanup_divergence_magnetic, cleanupDivergenceMagnetic) and Fourier filtering of newE, newB
(pic_filter_fields, filterFields) on the GPU:
- whole steps with each of them after every push, against the compiled reference's golden vectors
  (tests/golden/cleanup_ref.npz, filter_ref.npz, cleanup_eb_ref.npz), every step's E, B, particles and
  energies;
- against the CPU restatement (oracle/cleanup_oracle.py) on random states, and Gauss's law
  div E = 4 pi rho after the cleanup.
The cleanup's GMRES stops at the reference's Givens estimate 1e-8 / (n0 n1 n2), so E after it agrees
with an exact solve to about that tolerance; the filter is a fixed linear map, FP64 to rounding.
The magnetic cleanup solves to a Givens estimate of 1e-15 and leaves div B at rounding. The
reference's BiCGStab (to 1e-8) does the same in steps 1-4 of the mini case (div B of its B about
5e-17 relative) but leaves 6.3e-12 in step 5, where its right part is rounding noise: there its B
differs from a solenoidal one by a gradient of that size, and the bar for B in that step is ten
times the reference's own relative div B, with div B of the GPU's B checked at rounding.
"""
import dataclasses
import os

import numpy as np
import pytest

from conftest import GoldenState, relerr
from oracle import cleanup_oracle as co
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# measured worst over the 5 steps: cleanup E 1.1e-13, filter E 4.6e-15 (particles, B, energies below)
BARS = {"cleanup_ref.npz": 1e-12, "filter_ref.npz": 1e-13, "cleanup_eb_ref.npz": 1e-12}
UNIT = (1.0, 1.0, 1.0)


def rel_div_b(B):
    o = slice(1, -2)  # the owned nodes
    return np.linalg.norm(co.div_b(B, UNIT)[o, o, o]) / np.linalg.norm(B)


@pytest.mark.parametrize("fixture", ["cleanup_ref.npz", "filter_ref.npz", "cleanup_eb_ref.npz"])
def test_steps_with_cleanup_or_filter_match_reference(fixture):
    d = np.load(os.path.join(HERE, "golden", fixture))
    w = wl.mini()
    if fixture == "cleanup_ref.npz":
        w = dataclasses.replace(w, divergence_cleanup=1, name="mini-cleanup")
    elif fixture == "cleanup_eb_ref.npz":
        w = dataclasses.replace(w, divergence_cleanup=3, name="mini-cleanup-eb")
    else:
        w = dataclasses.replace(w, filter_cut_wave_number=2, name="mini-filter")
    print("case", w.name, fixture, flush=True)
    sim = Simulation(w)
    sim.load_state(GoldenState(d))
    sim.initialize()
    worst = {}
    for step in range(1, w.steps + 1):
        sim.step()
        g = sim.particles()
        gi = np.argsort(g["ids"])
        ref = d["s%d/particles" % step]
        ref = ref[np.argsort(ref[:, 0])]
        assert np.array_equal(g["ids"][gi].astype(np.float64), ref[:, 0])
        cur = {"x": relerr(g["x"][gi], ref[:, 1:4]), "p": relerr(g["p"][gi], ref[:, 4:7]),
               "E": relerr(sim.array("E"), d["s%d/E" % step]), "B": relerr(sim.array("B"), d["s%d/B" % step]),
               "energy": relerr(sim.updateEnergy()[:4], d["s%d/energy" % step][:4])}
        if w.divergence_cleanup & 2:
            # B against the reference within its own solver residual (module docstring); the
            # field energy follows B, so the energies get the same bar
            ref_div = rel_div_b(d["s%d/B" % step])
            assert rel_div_b(sim.array("B")) <= 1e-14, step
            slack = max(BARS[fixture], 10 * ref_div)
            print("   step %d: B %.1e, energy %.1e, reference's own div B %.1e" % (step, cur["B"], cur["energy"], ref_div))
            assert cur["B"] <= slack and cur["energy"] <= slack, (step, cur, ref_div)
            cur["B"] = cur["energy"] = 0.0
        for k, v in cur.items():
            worst[k] = max(worst.get(k, 0.0), v)
    print("%s, GPU vs reference over %d steps (worst):" % (w.name, w.steps), {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= BARS[fixture], (k, v)
    sim.close()


def test_cleanup_against_exact_solve_and_gauss_law():
    """Odd cell counts (no modes at k = pi, which the averaged gradient and divergence cannot
    reach): after the cleanup div newE = 4 pi rho to the GMRES tolerance, and newE equals the
    oracle's exact Fourier solve to that tolerance."""
    w = wl.mini(n=(9, 7, 5), ppc=6, seed=12)
    st = wl.make_initial_state(w, perturb=0.1)
    print("case cleanup", w.n, flush=True)
    sim = Simulation(w)
    sim.load_state(st)
    sim.initialize()
    rho = sim.array("density")[..., 5 * len(w.species)]
    E0 = sim.array("Enew")
    info = sim.cleanupDivergence()
    E1 = sim.array("Enew")
    d = (w.dx,) * 3
    want, _ = co.cleanup_divergence(E0, rho, d)
    o = slice(2, -2)
    gauss = relerr(co.div_e(E1, d)[o, o, o], 4 * np.pi * rho[o, o, o])
    before = relerr(co.div_e(E0, d)[o, o, o], 4 * np.pi * rho[o, o, o])
    err = relerr(E1, want)
    print("cleanup: %d GMRES steps, Gauss residual %.1e (before %.1e), vs exact solve %.1e" % (
        info["iterations"], gauss, before, err))
    assert before > 0.1
    assert gauss <= 1e-9  # measured 3.3e-11 after 45 GMRES steps
    assert err <= 1e-9  # measured 3.1e-11
    sim.close()


def test_magnetic_cleanup_against_exact_solve():
    w = wl.mini(n=(9, 7, 5), ppc=2, seed=15)
    st = wl.make_initial_state(w)
    rng = np.random.default_rng(6)
    st = dataclasses.replace(st, B=wl.periodic_fill_cells(rng.normal(size=w.n + (3,))))
    print("case magnetic cleanup", w.n, flush=True)
    sim = Simulation(w)
    sim.load_state(st)
    info = sim.cleanupDivergenceMagnetic()
    B1 = sim.array("Bnew")
    d = (w.dx,) * 3
    want, _ = co.cleanup_divergence_magnetic(st.B, d)
    o = slice(1, -2)
    before = np.linalg.norm(co.div_b(st.B, d)[o, o, o]) / np.linalg.norm(st.B)
    after = np.linalg.norm(co.div_b(B1, d)[o, o, o]) / np.linalg.norm(B1)
    err = relerr(B1, want)
    print("magnetic cleanup: %d GMRES steps, div B %.1e -> %.1e, vs exact solve %.1e" % (info["iterations"], before, after, err))
    assert before > 0.1 and after <= 1e-14 and err <= 1e-13  # measured 4.0e-16, 1.1e-15
    sim.close()


def test_filter_against_oracle_on_random_fields():
    w = wl.mini(n=(12, 10, 8), ppc=2, seed=14)
    st = wl.make_initial_state(w)
    rng = np.random.default_rng(3)
    st = dataclasses.replace(st, E=wl.periodic_fill_nodes(rng.normal(size=w.n + (3,))),
                             B=wl.periodic_fill_cells(rng.normal(size=w.n + (3,))))
    print("case filter", w.n, flush=True)
    sim = Simulation(w)
    sim.load_state(st)
    for cut in (2, 3):
        sim.load_state(st)
        sim.filterFields(cut)
        Ew, Bw = co.filter_fields(st.E, st.B, cut, (w.dx,) * 3)
        eE, eB = relerr(sim.array("Enew"), Ew), relerr(sim.array("Bnew"), Bw)
        print("filter cut %d: E %.1e, B %.1e" % (cut, eE, eB))
        assert eE <= 1e-13 and eB <= 1e-13
    sim.close()
