"""The HIP explicit (BUNEMAN / Tristan) step (PIC_SCHEME_EXPLICIT, pic_explicit_step) against
the compiled reference's own explicit step (tests/golden/explicit_ref.npz, made by
oracle/ref_harness_explicit.cpp; the CPU restatement oracle/explicit_oracle.cpp matches it
bit for bit, tests/test_explicit_oracle.py).

The case: 6x5x5 cells, 4 ppc per species, dt = 0.4, fast electrons (u ~ N(0, 0.5) per axis)
so that many zig-zag paths cross cells and some particles leave through every face, perturbed
Yee fields, 6 steps. Particles (matched by id), the Yee E and B after every step, the Yee
current of the first two steps and the energies and total momentum of updateEnergy (from the
Yee components) must agree to 1e-12 normwise; the fixed-point current makes two runs bitwise
identical.

A second case checks the paths that the small case cannot reach against the CPU restatement
(oracle/explicit_oracle.cpp): at 8x8x24 cells with 32 ppc per species, a workgroup's 256
particles start in one (i, j) row of about 8 cells, so the push sums the current of most
workgroups in its LDS box, and long runs of lanes share a cell in the wave's run sums.
"""
import dataclasses
import os

import numpy as np
import pytest

from conftest import relerr
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation
from oracle.oracle import OracleSim
from test_explicit_oracle import explicit_workload

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def open_explicit(d, w):
    N = w.N
    st = wl.InitialState(d["in/species"].astype(np.int32), d["in/ids"].astype(np.int32), d["in/x"], d["in/p"],
                         np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3)), np.zeros(N + (3,)))
    sim = Simulation(w, precision="fp64", scheme="explicit")
    sim.load_state(st)
    sim.load_yee({name: d["in/" + name] for name in wl.YEE_NAMES})
    return sim


def by_id(ids, x, p):
    o = np.argsort(ids)
    return ids[o], x[o], p[o]


def test_explicit_step_matches_reference():
    d = np.load(os.path.join(HERE, "golden", "explicit_ref.npz"))
    w = explicit_workload()
    print("case", w.name, "particles", len(d["in/ids"]), flush=True)
    sim = open_explicit(d, w)
    worst = {}
    for name in wl.YEE_NAMES:
        worst[name + "0"] = relerr(sim.yee(name), d["s0/" + name])

    def energy_errors(step):
        # updateEnergy with the Yee components (the explicit scheme's getBuneman* fields)
        e, want = np.asarray(sim.updateEnergy()), d["s%d/energy" % step]
        worst["energy"] = max(worst.get("energy", 0.0), relerr(e[:4], want[:4]))
        worst["momentum"] = max(worst.get("momentum", 0.0), relerr(e[4:], want[4:]))

    energy_errors(0)
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        energy_errors(step)
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        ref = d["s%d/particles" % step]
        ref = ref[np.argsort(ref[:, 0])]
        assert np.array_equal(ids.astype(np.float64), ref[:, 0]), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, ref[:, 1:4]))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, ref[:, 4:7]))
        for name in wl.YEE_NAMES:
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), d["s%d/%s" % (step, name)]))
        if step <= 2:
            for name in ("Jx", "Jy", "Jz"):
                worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), d["s%d/%s" % (step, name)]))
    print("explicit step vs reference:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_step_with_current_smoothing_matches_reference():
    """8 passes of smoothBunemanEfieldGeneral (s = 0.5, the shipped input's setting) on the current
    every step (tests/golden/explicit_smooth_ref.npz, from the compiled reference): the smoothing
    changes J by 99.7 % and E after 6 steps by 96 % against the unsmoothed run."""
    d = np.load(os.path.join(HERE, "golden", "explicit_smooth_ref.npz"))
    w = dataclasses.replace(explicit_workload(), smoothing_count=8, smoothing_parameter=0.5, name="explicit-smooth")
    print("case", w.name, flush=True)
    sim = open_explicit(d, w)
    worst = {}
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        ref = d["s%d/particles" % step]
        ref = ref[np.argsort(ref[:, 0])]
        assert np.array_equal(ids.astype(np.float64), ref[:, 0]), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, ref[:, 1:4]))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, ref[:, 4:7]))
        for name in wl.YEE_NAMES + (("Jx", "Jy", "Jz") if step <= 2 else ()):
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), d["s%d/%s" % (step, name)]))
        e = np.asarray(sim.updateEnergy())
        worst["energy"] = max(worst.get("energy", 0.0), relerr(e[:4], d["s%d/energy" % step][:4]))
    print("explicit step with smoothing vs reference:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def compare_steps(sim, d, w, worst):
    """Steps sim through w.steps and compares every step with the golden d: particles by id,
    the Yee fields, the current of the first two steps and updateEnergy's components."""
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        ref = d["s%d/particles" % step]
        ref = ref[np.argsort(ref[:, 0])]
        assert np.array_equal(ids.astype(np.float64), ref[:, 0]), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, ref[:, 1:4]))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, ref[:, 4:7]))
        for name in wl.YEE_NAMES + (("Jx", "Jy", "Jz") if step <= 2 else ()):
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), d["s%d/%s" % (step, name)]))
        e = np.asarray(sim.updateEnergy())
        worst["energy"] = max(worst.get("energy", 0.0), relerr(e[:4], d["s%d/energy" % step][:4]))
    return worst


def explicit_wall_workload(inject):
    return wl.shock(n=(28, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=6, seed=23,
                    name="explicit-wall", inject=inject)


def open_explicit_wall(d, w):
    N = w.N
    st = wl.InitialState(d["in/species"].astype(np.int32), d["in/ids"].astype(np.int32), d["in/x"], d["in/p"],
                         np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3)), np.zeros(N + (3,)))
    sim = Simulation(w, precision="fp64", scheme="explicit")
    sim.load_yee({name: d["in/" + name] for name in wl.YEE_NAMES})
    sim.load_state(st)
    return sim


def test_explicit_wall_matches_reference():
    """The wall (tests/golden/explicit_wall_ref.npz): the Tristan mover reflects at xgrid[20]
    (reflectingWallPoint) and starts the reflected path's current from the mirrored previous
    position; particles past node N - 2 leave through the free right end (their current still
    counted); the fields' conductor at node 2 zeroes E_y, E_z there, the free end holds E0, and the
    x ghost planes follow the one-rank non-periodic exchanges (zero or plane 2 per component, no
    current folds along x). Particles, Yee fields, current and energies within 1e-12 over 6 steps
    with about 45 escapes per step."""
    d = np.load(os.path.join(HERE, "golden", "explicit_wall_ref.npz"))
    w = explicit_wall_workload(False)
    print("case", w.name, "particles", len(d["in/ids"]), flush=True)
    sim = open_explicit_wall(d, w)
    worst = {}
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        ref = d["s%d/particles" % step]
        ref = ref[np.argsort(ref[:, 0])]
        assert np.array_equal(ids.astype(np.float64), ref[:, 0]), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, ref[:, 1:4]))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, ref[:, 4:7]))
        for name in wl.YEE_NAMES + (("Jx", "Jy", "Jz") if step <= 2 else ()):
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), d["s%d/%s" % (step, name)]))
        e = np.asarray(sim.updateEnergy())
        worst["energy"] = max(worst.get("energy", 0.0), relerr(e[:4], d["s%d/energy" % step][:4]))
    print("explicit wall vs reference:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_injection_matches_reference():
    """Injection at the right end and replacement of the escapes (explicit_inject_ref.npz, the
    reference's own injectNewParticles and addToPreserveChargeGlobal): every step adds as many
    particles as the reference; in the first step, before the new particles' current feeds back,
    the kept particles and the fields match within 1e-12 and the new particles' x positions are
    the reference's bit for bit (their momenta follow the reference's distribution, not its
    rand() samples)."""
    d = np.load(os.path.join(HERE, "golden", "explicit_inject_ref.npz"))
    w = explicit_wall_workload(True)
    n0 = len(d["in/ids"])
    print("case", w.name + "-inject", "particles", n0, flush=True)
    sim = open_explicit_wall(d, w)
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        p = sim.particles()
        ref = d["s%d/particles" % step]
        assert len(p["ids"]) == len(ref), (step, len(p["ids"]), len(ref))
        if step == 1:
            old = p["ids"] < n0
            ids, x, mom = by_id(p["ids"][old], p["x"][old], p["p"][old])
            r_old = ref[ref[:, 0] < n0]
            r_old = r_old[np.argsort(r_old[:, 0])]
            assert np.array_equal(ids.astype(np.float64), r_old[:, 0])
            assert relerr(x, r_old[:, 1:4]) <= 1e-12 and relerr(mom, r_old[:, 4:7]) <= 1e-12
            assert np.array_equal(np.sort(p["x"][~old][:, 0]), np.sort(ref[ref[:, 0] >= n0][:, 1]))
            for name in wl.YEE_NAMES:
                assert relerr(sim.yee(name), d["s1/" + name]) <= 1e-12, name
    print("particles per step:", [len(d["s%d/particles" % s]) for s in range(w.steps + 1)])
    sim.close()


def test_explicit_step_deterministic_bitwise():
    d = np.load(os.path.join(HERE, "golden", "explicit_ref.npz"))
    w = explicit_workload()
    print("case", w.name, "two runs of 3 steps", flush=True)
    out = []
    for _ in range(2):
        sim = open_explicit(d, w)
        for _ in range(3):
            sim.explicit_step()
        p = sim.particles()
        out.append([sim.yee(n) for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz")] + [p["x"], p["p"]])
        sim.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def box_case(n=(8, 8, 24), ppc=32, steps=3):
    w = wl.Workload("explicit-%dx%dx%d-ppc%d" % (n + (ppc,)), n, 1.0, 0.4, steps, wl.electron_proton(ppc, 0.01, 100.0),
                    b0=(wl.b_for_gyro(0.5), 0.0, 0.0), seed=5)
    st = wl.make_initial_state(w)
    rng = np.random.default_rng(31)
    e = st.species == 0
    st.p[e] = w.species[0].mass * rng.normal(0.0, 0.3, (int(e.sum()), 3))
    return w, st, wl.yee_initial_fields(w, perturb=0.05)


@pytest.mark.parametrize("n,ppc,steps", [((8, 8, 24), 32, 3), ((64, 64, 64), 4, 2)])
def test_explicit_step_matches_oracle_with_workgroup_boxes(n, ppc, steps):
    # the second case: 2.1 M particles on 68^3 ghosted cells, rows of 64 cells holding about
    # 256 particles per species, so many workgroups straddle two rows
    w, st, yee = box_case(n, ppc, steps)
    print("case", w.name, "particles", len(st.ids), flush=True)
    sim = Simulation(w, precision="fp64", scheme="explicit")
    sim.load_state(st)
    sim.load_yee(yee)
    o = OracleSim(w, st)
    o.yee_init()
    for name in wl.YEE_NAMES:
        o.set_yee(name, yee[name])
    o.yee_exchange()
    worst = {}
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        o.explicit_step()
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        rids, rx, rp = by_id(*(o.particles()[k] for k in ("ids", "x", "p")))
        assert np.array_equal(ids, rids), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, rx))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, rp))
        for name in wl.YEE_NAMES + ("Jx", "Jy", "Jz"):
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), o.yee(name)))
    print("explicit step vs oracle:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_cold_plasma_matches_oracle():
    """A cold plasma (kT/mc^2 = 1e-4: electrons at u_th = 0.01, protons at 0.001) at 64 ppc per
    species. The fixed-point quantum of the current follows the particles' actual speeds (their
    largest |u| plus what the largest |E| can add in a step), not the speed of light, so the small
    contributions of slow particles keep FP64 accuracy: J, E, B, x and p within 1e-12 of the
    oracle's FP64 sums."""
    n, ppc, steps = (16, 16, 16), 64, 3
    w = wl.Workload("explicit-cold-%dx%dx%d-ppc%d" % (n + (ppc,)), n, 1.0, 0.4, steps, wl.electron_proton(ppc, 1e-4, 100.0),
                    b0=(wl.b_for_gyro(0.5), 0.0, 0.0), seed=9)
    st = wl.make_initial_state(w)
    yee = wl.yee_initial_fields(w, perturb=0.01)
    print("case", w.name, "particles", len(st.ids), flush=True)
    sim = Simulation(w, precision="fp64", scheme="explicit")
    sim.load_state(st)
    sim.load_yee(yee)
    o = OracleSim(w, st)
    o.yee_init()
    for name in wl.YEE_NAMES:
        o.set_yee(name, yee[name])
    o.yee_exchange()
    worst = {}
    for step in range(1, steps + 1):
        sim.explicit_step()
        o.explicit_step()
        p = sim.particles()
        ids, x, mom = by_id(p["ids"], p["x"], p["p"])
        rids, rx, rp = by_id(*(o.particles()[k] for k in ("ids", "x", "p")))
        assert np.array_equal(ids, rids), step
        worst["x"] = max(worst.get("x", 0.0), relerr(x, rx))
        worst["p"] = max(worst.get("p", 0.0), relerr(mom, rp))
        for name in wl.YEE_NAMES + ("Jx", "Jy", "Jz"):
            worst[name] = max(worst.get(name, 0.0), relerr(sim.yee(name), o.yee(name)))
    print("cold plasma, explicit step vs oracle:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_step_without_particles():
    # no particles: the Yee leapfrog alone (a vacuum), against the oracle; the current stays 0
    w, st, yee = box_case((6, 5, 7), 1, 3)
    empty = wl.InitialState(st.species[:0], st.ids[:0], st.x[:0], st.p[:0], st.E, st.B)
    print("case", w.name, "without particles", flush=True)
    sim = Simulation(w, precision="fp64", scheme="explicit")
    sim.load_state(empty)
    sim.load_yee(yee)
    o = OracleSim(w, empty)
    o.yee_init()
    for name in wl.YEE_NAMES:
        o.set_yee(name, yee[name])
    o.yee_exchange()
    for step in range(w.steps):
        sim.explicit_step()
        o.explicit_step()
        assert sim.num_particles() == 0
        for name in wl.YEE_NAMES:
            err = relerr(sim.yee(name), o.yee(name))
            assert err <= 1e-13, (step, name, err)
        for name in ("Jx", "Jy", "Jz"):
            assert not np.any(sim.yee(name)), name
    sim.close()


def explicit2d_workload():
    return dataclasses.replace(wl.mini(n=(8, 6, 1), ppc=4), dt=0.4, steps=6, name="explicit-2d")


def explicit2d_shock_workload(inject):
    w = wl.shock(n=(28, 6, 1), ppc=10, gamma0=1.5, sigma=0.03, theta=1.15e-3, mass_ratio=25.0, dt=0.5, steps=6, seed=37,
                 name="explicit-2d-shock", inject=inject)
    return dataclasses.replace(w, smoothing_count=8, smoothing_parameter=0.5)


def test_explicit_2d_matches_reference():
    """The 2-D XY grid (8x6x1 cells; tests/golden/explicit2d_ref.npz, the reference with
    znumberGeneral = 1): one plane along z without ghost layers, no z derivatives in the curls,
    the TWO_D_XY gather and zig-zag current (weight 1 on plane 0, the z relay at the midpoint),
    plane 1 of the two-plane arrays copied from plane 0 by every exchange, z unwrapped. Particles,
    Yee fields, current and energies within 1e-12 over 6 steps."""
    d = np.load(os.path.join(HERE, "golden", "explicit2d_ref.npz"))
    w = explicit2d_workload()
    print("case", w.name, "particles", len(d["in/ids"]), flush=True)
    sim = open_explicit(d, w)
    worst = {name + "0": relerr(sim.yee(name), d["s0/" + name]) for name in wl.YEE_NAMES}
    compare_steps(sim, d, w, worst)
    print("explicit 2-D vs reference:", {k: "%.1e" % v for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_2d_shock_matches_reference():
    """The shipped input's kind of run on a 2-D XY grid (tests/golden/explicit2d_shock_ref.npz:
    28x6x1 cells, a gamma 1.5 flow, B along z, mass ratio 25, 10 ppc per species, the particles'
    wall at xgrid[20], escapes through the free right end, 8 smoothing passes of s = 0.5 every
    step, 2-D smoothing kernel without z neighbours): within 1e-12 over 6 steps."""
    d = np.load(os.path.join(HERE, "golden", "explicit2d_shock_ref.npz"))
    w = explicit2d_shock_workload(False)
    print("case", w.name, "particles", len(d["in/ids"]), flush=True)
    sim = open_explicit_wall(d, w)
    worst = compare_steps(sim, d, w, {})
    print("explicit 2-D shock vs reference:", {k: "%.1e" % v for k, v in worst.items()},
          "particles", [len(d["s%d/particles" % s]) for s in range(w.steps + 1)])
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    sim.close()


def test_explicit_2d_injection_matches_reference():
    """The same with injection and replacement (tests/golden/explicit2d_inject_ref.npz): the
    reference's particle count every step; in the first step the kept particles and the fields
    within 1e-12 and the new particles' x positions bit for bit; the new particles' z in the one
    cell along z (ygrid[j] + dy, zgrid[0] + dz), the replacements' z their escaped particle's."""
    d = np.load(os.path.join(HERE, "golden", "explicit2d_inject_ref.npz"))
    w = explicit2d_shock_workload(True)
    n0 = len(d["in/ids"])
    print("case", w.name + "-inject", "particles", n0, flush=True)
    sim = open_explicit_wall(d, w)
    for step in range(1, w.steps + 1):
        sim.explicit_step()
        p = sim.particles()
        ref = d["s%d/particles" % step]
        assert len(p["ids"]) == len(ref), (step, len(p["ids"]), len(ref))
        if step == 1:
            old = p["ids"] < n0
            ids, x, mom = by_id(p["ids"][old], p["x"][old], p["p"][old])
            r_old = ref[ref[:, 0] < n0]
            r_old = r_old[np.argsort(r_old[:, 0])]
            assert np.array_equal(ids.astype(np.float64), r_old[:, 0])
            assert relerr(x, r_old[:, 1:4]) <= 1e-12 and relerr(mom, r_old[:, 4:7]) <= 1e-12
            assert np.array_equal(np.sort(p["x"][~old][:, 0]), np.sort(ref[ref[:, 0] >= n0][:, 1]))
            for name in wl.YEE_NAMES:
                assert relerr(sim.yee(name), d["s1/" + name]) <= 1e-12, name
            # the slabs injected in this step have not moved yet: z in the one cell along z
            z0 = w.origin()[2]
            inj = ~old & (p["x"][:, 0] < w.origin()[0] + (w.N[0] - 2) * w.dx)
            assert inj.sum() > 0 and np.all((p["x"][inj, 2] >= z0) & (p["x"][inj, 2] < z0 + w.dx))
    print("particles per step:", [len(d["s%d/particles" % s]) for s in range(w.steps + 1)])
    sim.close()


def test_2d_grid_needs_the_explicit_scheme():
    """A one-cell z is built for the explicit scheme only: the implicit scheme refuses it."""
    from picpp_amd._lib import PicError
    w = dataclasses.replace(wl.mini(n=(8, 6, 1), ppc=1), name="implicit-2d")
    with pytest.raises(PicError):
        Simulation(w, precision="fp64")
