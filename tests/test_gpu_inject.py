"""Injection at the free right end and replacement of escaped particles on the HIP path
(pic_push with inject on), against the CPU restatement (oracle/inject_oracle.py, pinned by
tests/test_inject_oracle.py) and the compiled reference's own injection
(tests/golden/inject_ref.npz).

The case: the wall case with the right half of the particles moving along +x (some leave
through the free right end) and a gamma 2 flow injected at the right end, 12x5x5 cells, 4 ppc
per species, dt = 0.5, 5 steps. Every step must
- inject and replace exactly as many particles per species as the reference does;
- produce the oracle's new particles: FP64 ids, species and positions exact, momenta within
  1e-13 (the two sides' log, sin and cos may differ in the last bit); a replacement's y and z
  lie in the box (they come from the escaped particle, which neither side exposes);
- in the first step, before injected particles feed back into the fields, move every other
  particle as the reference does (1e-12).
The FP32 form injects the same particles, stored as cell + FP32 offset and FP32 u: positions
within 1e-6 cells, momenta within 1e-6.
The same holds with hot injected electrons (tests/golden/inject_hot_ref.npz: kT / m_e c^2 = 0.5,
momenta from the tabulated Maxwell-Juttner inverse; inject_hotrun_ref.npz: the run's
kT / m_e c^2 = 3, the rejection sampler): per-species counts as the reference's, the oracle's
particles, and the first-step push of the others.
"""
import os

import numpy as np
import pytest

from conftest import GoldenState, relerr
from oracle import inject_oracle as io
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def inject_workload():
    return wl.shock(n=(12, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=5, seed=19, name="inject")


def test_philox_matches_oracle():
    import torch
    assert torch.cuda.is_available()
    print("case: philox counters", flush=True)
    sim = Simulation(wl.mini(n=(6, 5, 5), ppc=1, seed=1))
    n = 64
    out = torch.zeros(4 * n, dtype=torch.int64, device="cuda")
    for ctr, key in (((0, 0, 0, 0), (0, 0)), ((3, (1 << 32) | 17, 5, 0), (19, 0)), ((2 ** 63, 2 ** 64 - 40, 7, 9), (2 ** 64 - 1, 5))):
        sim.philox(ctr, key, n, out.data_ptr())
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint64).reshape(n, 4)
        for i in range(n):
            c = list(ctr)
            c[1] = (c[1] + i) % (1 << 64)
            assert [int(v) for v in got[i]] == io.philox4x64_10(c, key), (ctr, i)
    sim.close()


def hot_workload(run_kT=0.0):
    """Hot injected electrons (kT / m_e c^2 = 0.5: the tabulated Maxwell-Juttner sampler; with
    the run's kT / m_e c^2 = 3 the rejection sampler), as inject_hot_ref.npz and
    inject_hotrun_ref.npz were made."""
    import dataclasses
    sp = wl.electron_proton(4, 0.5, 100.0)
    for s in sp:
        s.drift_gamma, s.drift_axis, s.drift_sign = 2.0, 0, -1
    return dataclasses.replace(inject_workload(), name="inject-hot", run_kT=run_kT, species=sp)


def run_case(precision, fixture="inject_ref.npz", w=None):
    d = np.load(os.path.join(HERE, "golden", fixture))
    w = w or inject_workload()
    st = GoldenState(d)
    print("case", w.name, precision, "particles", len(st.species), flush=True)
    sim = Simulation(w, precision=precision)
    sim.load_state(st)
    sim.initialize()
    inj = io.Injector(w, next_id=int(st.ids.max()) + 1)
    xr = inj.right_edge()
    x0 = w.origin()
    ptype = np.array([sp.ptype for sp in w.species])
    errs = {}
    for step in range(1, w.steps + 1):
        before = sim.particles()
        sim.step()
        after = sim.particles()
        # escapes in the order of their index before the push: the replacement order
        kept = set(after["ids"].tolist())
        esc = [(int(s), np.nan, np.nan) for i, s in zip(before["ids"], before["species"]) if int(i) not in kept]
        want = inj.step_particles(esc)
        ref = d["s%d/particles" % step]
        rtypes = d["s%d/ptype" % step].astype(int)
        assert len(after["ids"]) == len(ref), step
        assert np.array_equal(np.bincount(ptype[after["species"]], minlength=3), np.bincount(rtypes, minlength=3)), step
        new = after["ids"] >= want["ids"][0] if len(want["ids"]) else np.zeros(len(after["ids"]), bool)
        order = np.argsort(after["ids"][new])
        g_ids = after["ids"][new][order]
        g_sp = after["species"][new][order]
        g_x = after["x"][new][order]
        g_p = after["p"][new][order]
        assert np.array_equal(g_ids, want["ids"]), step
        assert np.array_equal(g_sp, want["species"]), step
        inj_m = want["kind"] == 0
        rep_m = want["kind"] == 1
        if precision == "fp64":
            assert np.array_equal(g_x[inj_m], want["x"][inj_m]), step
            assert np.all(g_x[rep_m, 0] == xr), step
            ep = relerr(g_p, want["p"])
            assert ep <= 1e-13, (step, ep)
        else:
            dx_cells = np.abs(g_x[inj_m] - want["x"][inj_m]) / w.dx
            assert dx_cells.max() <= 1e-6, (step, dx_cells.max())
            assert np.all(np.abs(g_x[rep_m, 0] - xr) <= 1e-6 * w.dx), step
            ep = relerr(g_p, want["p"])
            assert ep <= 1e-6, (step, ep)
        errs["p"] = max(errs.get("p", 0.0), ep)
        for a in (1, 2):
            lo, hi = x0[a] + 2 * w.dx, x0[a] + (w.n[a] + 2) * w.dx
            assert np.all((g_x[rep_m, a] >= lo) & (g_x[rep_m, a] <= hi)), step
        if step == 1:  # the particles that were there before, against the reference's push
            old = ~new
            gi = np.argsort(after["ids"][old])
            r = ref[np.argsort(ref[:, 0])]
            r = r[r[:, 0] < want["ids"][0]]
            assert np.array_equal(after["ids"][old][gi].astype(np.float64), r[:, 0])
            ex = relerr(after["x"][old][gi], r[:, 1:4])
            ep1 = relerr(after["p"][old][gi], r[:, 4:7])
            errs["x_step1"], errs["p_step1"] = ex, ep1
            tol = 1e-12 if precision == "fp64" else 2e-5
            assert ex <= tol and ep1 <= tol, (ex, ep1)
        errs.setdefault("injected", 0)
        errs["injected"] += int(inj_m.sum())
        errs.setdefault("replaced", 0)
        errs["replaced"] += int(rep_m.sum())
    print(precision, "injection vs oracle and reference:", errs, "particles", len(after["ids"]))
    sim.close()


def test_injection_fp64_matches_oracle_and_reference():
    run_case("fp64")


def test_injection_fp32_matches_oracle_and_reference():
    run_case("fp32")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("fixture,run_kT", [("inject_hot_ref.npz", 0.0), ("inject_hotrun_ref.npz", 3.0)])
def test_hot_injection_matches_oracle_and_reference(precision, fixture, run_kT):
    run_case(precision, fixture, hot_workload(run_kT))
