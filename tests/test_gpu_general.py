"""general.dat rows on the GPU (pic_general, pic_write_general; outputGeneral,
PIC++/output.cpp:3215-3250) against the compiled reference's own rows
(tests/golden/general_ref.npz: the perturbed mini case, periodic; general_wall_ref.npz: the wall
case with reflections at the wall and escapes through the free right end; 5 steps each, made by
oracle/ref_harness.cpp with the loop's updateTheoreticalEnergy and updateParameters):
- every number of every row (energies and their components, momenta, the theoretical energy and
  momentum, the largest |E| and |B|, the rms of E, the particle count) within 1e-12 of the
  reference, relative to the row's total energy or largest momentum;
- the rows written by pic_write_general against the reference's text, field by field;
- the injection's share of the theoretical totals: in the first step of the injection case the
  totals exceed those of the same step without injection by exactly the energy and momentum of
  the new particles (the injected momenta are not the reference's rand() samples, so that case
  has no reference row after the first);
- a checkpoint carries the totals: a resumed run repeats the uninterrupted rows bit for bit.
"""
import os

import numpy as np
import pytest

from conftest import GoldenState
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ENERGY = [0, 1, 2, 3, 7, 17, 18, 19, 20, 21, 22]
MOMENTUM = [4, 5, 6, 8, 9, 10, 23, 24, 25, 26, 27, 28]
FIELDS = [11, 12, 14, 15, 16]


def wall_workload(inject=False):
    return wl.shock(n=(12, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=5, seed=19,
                    name="wall-general", inject=inject)


def workload_for(fixture):
    return wl.mini() if fixture == "general_ref.npz" else wall_workload()


def row_errors(got, want):
    e_scale = abs(want[3])
    p_scale = np.max(np.abs(want[MOMENTUM]))
    return {"energy": float(np.max(np.abs(got[ENERGY] - want[ENERGY])) / e_scale),
            "momentum": float(np.max(np.abs(got[MOMENTUM] - want[MOMENTUM])) / p_scale),
            "fields": float(np.max(np.abs(got[FIELDS] - want[FIELDS]) / np.abs(want[FIELDS])))}


@pytest.mark.parametrize("fixture", ["general_ref.npz", "general_wall_ref.npz"])
def test_general_rows_match_reference(fixture, tmp_path):
    d = np.load(os.path.join(HERE, "golden", fixture))
    w = workload_for(fixture)
    print("case", fixture, w.name, flush=True)
    sim = Simulation(w)
    sim.load_state(GoldenState(d))
    sim.initialize()
    out = str(tmp_path / "general.dat")
    worst = {}
    for step in range(w.steps + 1):
        if step > 0:
            sim.step()
        got = sim.general()
        want = d["s%d/general" % step]
        assert int(got[13]) == int(want[13]), (step, got[13], want[13])
        for k, v in row_errors(got, want).items():
            worst[k] = max(worst.get(k, 0.0), v)
        sim.outputGeneral(out, time=step * w.dt, current_iteration=step)
    print("general rows vs reference (worst over %d steps):" % w.steps, {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    # the text: tokens 3..33 hold the row's numbers (10 significant digits), the rest (iteration,
    # times, dt, shockWaveX and the derivative points) must be the same text
    lines = open(out).read().splitlines()
    ref = [str(t) for t in d["general_text"]]
    assert len(lines) == len(ref)
    token_row = dict(zip(range(3, 16), range(0, 13)))
    token_row.update({17: 13})
    token_row.update(dict(zip(range(19, 34), range(14, 29))))
    same = 0
    for step, (a, b) in enumerate(zip(lines, ref)):
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb) == 38
        same += a == b
        want = d["s%d/general" % step]
        scale = {r: abs(want[3]) for r in ENERGY}
        scale.update({r: np.max(np.abs(want[MOMENTUM])) for r in MOMENTUM})
        scale.update({r: abs(want[r]) for r in FIELDS + [13]})
        for t, (x, y) in enumerate(zip(ta, tb)):
            if t in token_row:  # the printed values: their 10 digits, or the row's scale
                assert abs(float(x) - float(y)) <= 1e-9 * abs(float(y)) + 1e-12 * scale[token_row[t]], (step, t, x, y)
            else:
                assert x == y, (step, t, x, y)
    print("text rows identical to the reference's: %d of %d" % (same, len(ref)))
    sim.close()


def test_injection_adds_the_new_particles_to_the_theoretical_totals():
    d = np.load(os.path.join(HERE, "golden", "inject_ref.npz"))
    st = GoldenState(d)
    rows = {}
    parts = {}
    for inject in (False, True):
        w = wall_workload(inject=inject)
        sim = Simulation(w)
        sim.load_state(st)
        sim.initialize()
        sim.step()
        rows[inject] = sim.general()
        parts[inject] = sim.particles()
        sim.close()
    m, q, wt, _ = wl.species_arrays(wall_workload())
    p = parts[True]
    new = p["ids"] >= len(st.ids)  # injected and replacement particles: ids from one past the largest
    assert new.sum() > 0
    s = p["species"][new]
    mom = p["p"][new]
    gm = np.sqrt(np.sum(mom * mom, axis=1) / m[s] ** 2 + 1.0)
    de = np.sum((gm - 1) * m[s] * wt[s])
    dp = np.sum(mom * wt[s][:, None], axis=0)
    err_e = abs((rows[True][7] - rows[False][7]) - de) / abs(rows[False][3])
    err_p = np.max(np.abs((rows[True][8:11] - rows[False][8:11]) - dp)) / np.max(np.abs(rows[False][8:11]))
    print("new particles: %d, theoretical energy share %.6e (error %.1e), momentum error %.1e" % (new.sum(), de, err_e, err_p))
    assert err_e <= 1e-12 and err_p <= 1e-12


def test_checkpoint_carries_the_theoretical_totals(tmp_path):
    d = np.load(os.path.join(HERE, "golden", "general_wall_ref.npz"))
    w = wall_workload()
    ck = str(tmp_path / "run.ckpt")
    a = Simulation(w)
    a.load_state(GoldenState(d))
    a.initialize()
    for _ in range(2):
        a.step()
    a.save_checkpoint(ck)
    for _ in range(3):
        a.step()
    want = a.general()
    a.close()
    b = Simulation(w)
    b.load_checkpoint(ck)
    for _ in range(3):
        b.step()
    got = b.general()
    b.close()
    assert np.array_equal(got, want), np.max(np.abs(got - want))
