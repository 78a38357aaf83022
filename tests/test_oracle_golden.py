"""The FP64 oracle against golden vectors written by the compiled reference.

tests/golden/mini_ref.npz and config1_ref.npz hold seeded inputs and the outputs of
the unmodified PIC++ sources (built by oracle/Makefile.ref, run by
tests/golden/make_golden.py). The oracle restates the reference operation by
operation, so the comparison bar here is bit-exact for every phase.
"""
import os

import numpy as np
import pytest

from conftest import GoldenState, relerr
from oracle.oracle import OracleSim
from picpp_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def by_id(ids, arr):
    return arr[np.argsort(ids)]


def test_mini_every_phase_bit_exact(mini_golden):
    d = mini_golden
    o = OracleSim(wl.mini(), GoldenState(d))
    o.initial_sequence()
    assert np.array_equal(o.array("E"), d["s0/E"])
    assert np.array_equal(o.array("rho"), d["s0/rho"])
    assert np.array_equal(o.energy(), d["s0/energy"])
    for s in range(1, 6):
        o.evaluate_rotation_tensor()
        if s <= 2:
            P = o.particles()
            ref = d["s%d/alpha" % s]
            assert np.array_equal(by_id(P["ids"], P["alpha"]), by_id(ref[:, 0], ref[:, 1:]))
        o.update_em_parameters()
        if s <= 2:
            for name in ("J", "mu", "rhohat", "Pi", "divPi"):
                assert np.array_equal(o.array(name), d["s%d/%s" % (s, name)]), name
            o.build_rhs()
            sl = (slice(2, -2),) * 3  # interior rows (ghost rows are identity rows with zero right part)
            assert np.array_equal(o.array("rhs").reshape(d["s%d/rhs" % s].shape)[sl], d["s%d/rhs" % s][sl])
        o.evaluate_electric_field()
        o.exchange_fields()
        o.evaluate_magnetic_field()
        o.exchange_fields()
        if s <= 2:
            for name in ("Etheta", "Enew", "Bnew"):
                assert np.array_equal(o.array(name), d["s%d/%s" % (s, name)]), name
        o.move_particles()
        o.update_density_parameters()
        o.exchange_fields()
        o.update_fields()
        o.exchange_fields()
        o.update_energy()
        P = o.particles()
        ref = d["s%d/particles" % s]
        # same particle order as the reference's vector after erase/exchange
        assert np.array_equal(P["ids"], ref[:, 0].astype(np.int32))
        assert np.array_equal(P["x"], ref[:, 1:4])
        assert np.array_equal(P["p"], ref[:, 4:7])
        assert np.array_equal(o.array("E"), d["s%d/E" % s])
        assert np.array_equal(o.array("B"), d["s%d/B" % s])
        assert np.array_equal(o.array("rho"), d["s%d/rho" % s])
        assert np.array_equal(o.energy(), d["s%d/energy" % s])
        if s <= 2:
            for t in range(2):
                for name in ("conc%d" % t, "energy%d" % t, "bulkv%d" % t):
                    assert np.array_equal(o.array(name), d["s%d/%s" % (s, name)]), name


def test_config1_fifty_steps():
    """The reference's CPU case (32x32x8, 20 ppc per species, 50 steps): energies each
    step, field norms and 1024 sampled particles at steps 1, 10 and 50."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "config1_ref.npz"))
    w = wl.config1()
    st = wl.make_initial_state(w)
    o = OracleSim(w, st)
    o.initial_sequence()
    sample = d["sample_ids"]
    for s in range(0, w.steps + 1):
        if s > 0:
            o.step(1)
        assert relerr(o.energy(), d["s%d/energy" % s]) <= 1e-13, s
        for f in ("E", "B", "rho"):
            a = o.array(f)
            got = np.array([np.sqrt(np.sum(a * a)), np.sum(a)])
            ref = d["s%d/%s_norm" % (s, f)]
            assert abs(got[0] - ref[0]) <= 1e-12 * max(ref[0], 1e-300), (s, f)
        if s in (1, 10, w.steps):
            P = o.particles()
            order = np.argsort(P["ids"])
            got = np.concatenate([P["ids"][order][sample, None], P["x"][order][sample], P["p"][order][sample]], 1)
            ref = d["s%d/particles_sample" % s]
            assert np.array_equal(got[:, 0], ref[:, 0])
            assert relerr(got[:, 1:4], ref[:, 1:4]) <= 1e-14
            assert relerr(got[:, 4:7], ref[:, 4:7]) <= 1e-12
