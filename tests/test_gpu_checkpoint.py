"""Checkpoints (pic_checkpoint_save / pic_checkpoint_load, the library's own lossless format)
and the reference's backup files (pic_write_backup, outputSimulationBackup).

- A run interrupted by a checkpoint and resumed in a fresh context repeats the uninterrupted
  run's bits: FP64 and FP32 implicit steps (mini case) and explicit steps (the explicit golden
  case). The FP32 load forms the post-push pass's cell moments again, so the first resumed step
  uses the same moments as the uninterrupted run.
- A checkpoint does not load into a context of another configuration; a truncated file and one
  with a damaged end are refused before anything is written, so the context keeps its state.
- The wall case with injection resumes bit for bit (the injection's random stream and counters
  travel with the checkpoint).
- The backup files of the mini case's initial state are byte for byte the compiled
  reference's (tests/golden/backup_ref.npz, written by oracle/ref_harness.cpp when given a
  path prefix).
"""
import dataclasses
import os

import numpy as np
import pytest

from picpp_amd import _lib
from picpp_amd import workloads as wl
from picpp_amd.simulation import Simulation
from test_gpu_explicit import open_explicit
from test_explicit_oracle import explicit_workload

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def state(sim, explicit=False):
    p = sim.particles()
    out = {"ids": p["ids"], "x": p["x"], "p": p["p"]}
    names = wl.YEE_NAMES if explicit else ("E", "B")
    for n in names:
        out[n] = sim.yee(n) if explicit else sim.array(n)
    return out


def assert_bitwise(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_checkpoint_resume_implicit(tmp_path, precision):
    w = wl.mini()
    st = wl.make_initial_state(w, perturb=0.05)
    print("case", w.name, precision, flush=True)
    a = Simulation(w, precision=precision)
    a.load_state(st)
    a.initialize()
    for _ in range(2):
        a.step()
    path = str(tmp_path / "run.ckpt")
    a.save_checkpoint(path)
    saved = state(a)
    for _ in range(3):
        a.step()
    b = Simulation(w, precision=precision)
    b.load_checkpoint(path)
    assert_bitwise(saved, state(b))
    for _ in range(3):
        b.step()
    assert_bitwise(state(a), state(b))
    a.close()
    b.close()


def test_checkpoint_resume_explicit(tmp_path):
    d = np.load(os.path.join(HERE, "golden", "explicit_ref.npz"))
    w = explicit_workload()
    print("case", w.name, flush=True)
    a = open_explicit(d, w)
    for _ in range(2):
        a.explicit_step()
    path = str(tmp_path / "run.ckpt")
    a.save_checkpoint(path)
    for _ in range(3):
        a.explicit_step()
    b = Simulation(w, precision="fp64", scheme="explicit")
    b.load_checkpoint(path)
    for _ in range(3):
        b.explicit_step()
    assert_bitwise(state(a, True), state(b, True))
    a.close()
    b.close()


def test_checkpoint_resume_explicit_2d_shock_with_injection(tmp_path):
    """The 2-D XY shock run (one cell along z: wall, injection, replacement, 8 smoothing passes;
    tests/golden/explicit2d_inject_ref.npz): saved after 2 steps, resumed in a fresh context,
    3 more steps give the uninterrupted run's bits, the injected particles included."""
    from test_gpu_explicit import explicit2d_shock_workload, open_explicit_wall
    d = np.load(os.path.join(HERE, "golden", "explicit2d_inject_ref.npz"))
    w = explicit2d_shock_workload(True)
    print("case", w.name, flush=True)
    a = open_explicit_wall(d, w)
    for _ in range(2):
        a.explicit_step()
    path = str(tmp_path / "run.ckpt")
    a.save_checkpoint(path)
    for _ in range(3):
        a.explicit_step()
    b = Simulation(w, precision="fp64", scheme="explicit")
    b.load_checkpoint(path)
    for _ in range(3):
        b.explicit_step()
    assert_bitwise(state(a, True), state(b, True))
    a.close()
    b.close()


def test_checkpoint_refuses_other_configuration_and_truncation(tmp_path):
    w = wl.mini()
    st = wl.make_initial_state(w, perturb=0.05)
    a = Simulation(w, precision="fp64")
    a.load_state(st)
    a.initialize()
    path = str(tmp_path / "run.ckpt")
    a.save_checkpoint(path)
    for other in (dataclasses.replace(w, dt=0.2), dataclasses.replace(w, n=(6, 5, 6))):
        b = Simulation(other, precision="fp64")
        with pytest.raises(_lib.PicError, match="checkpoint written with another"):
            b.load_checkpoint(path)
        b.close()
    c = Simulation(w, precision="fp32")
    with pytest.raises(_lib.PicError, match="precision"):
        c.load_checkpoint(path)
    c.close()
    cut = str(tmp_path / "cut.ckpt")
    with open(path, "rb") as f:
        data = f.read()
    with open(cut, "wb") as f:
        f.write(data[: len(data) - 100])
    bad_end = str(tmp_path / "bad_end.ckpt")  # the right length, a damaged end marker
    with open(bad_end, "wb") as f:
        f.write(data[:-1] + bytes([data[-1] ^ 0xFF]))
    # a refused file leaves the context as it was: a running context keeps its state
    d = Simulation(w, precision="fp64")
    d.load_state(wl.make_initial_state(dataclasses.replace(w, seed=w.seed + 1), perturb=0.05))
    d.initialize()
    d.step()
    before = state(d)
    for bad in (cut, bad_end):
        with pytest.raises(_lib.PicError, match="short read|truncated"):
            d.load_checkpoint(bad)
        assert_bitwise(state(d), before)
    d.step()  # and still steps
    d.close()
    a.close()


def test_checkpoint_resume_wall_with_injection(tmp_path):
    """The wall case with escapes, injection and replacement (the injection golden's state):
    the checkpoint carries the injection's stream step, slab count, injection length and next id,
    so a resumed run injects the same particles and repeats the uninterrupted run's bits."""
    d = np.load(os.path.join(HERE, "golden", "inject_ref.npz"))
    w = wl.shock(n=(12, 5, 5), ppc=4, gamma0=2.0, theta=0.01, mass_ratio=100.0, dt=0.5, steps=5, seed=19, name="inject")
    st = wl.InitialState(d["in/species"].astype(np.int32), d["in/ids"].astype(np.int32), d["in/x"], d["in/p"], d["in/E"],
                         d["in/B"])
    print("case", w.name, flush=True)
    a = Simulation(w, precision="fp64")
    a.load_state(st)
    a.initialize()
    for _ in range(2):
        a.step()
    path = str(tmp_path / "inject.ckpt")
    a.save_checkpoint(path)
    for _ in range(3):
        a.step()
    want = state(a)
    n_want = a.num_particles()
    a.close()
    b = Simulation(w, precision="fp64")
    b.load_checkpoint(path)
    for _ in range(3):
        b.step()
    assert b.num_particles() == n_want
    assert_bitwise(state(b), want)
    b.close()


def parse_general(text):
    """The general file's values in outputSimulationBackup's order (PIC++/output.cpp:3443-3540)."""
    v = text.split()
    it = iter(v)
    nxt = lambda: next(it)  # noqa: E731
    info = {"input_type": int(nxt())}
    for _ in range(4):
        nxt()  # x, y, z cell counts and the particle count: the context's own
    info["particles_per_bin"] = [int(nxt()) for _ in range(6)]
    info["concentration"] = [float(nxt()) for _ in range(6)]
    for key in ("temperature", "plasma_period", "scale_factor", "time", "max_time"):
        info[key] = float(nxt())
    info["current_iteration"] = int(nxt())
    info["max_iteration"] = int(nxt())
    info["size"] = [float(nxt()) for _ in range(3)]
    nxt(), nxt()  # theta, eta: the context's own
    info["debug_mode"] = int(nxt())
    nxt(), nxt()  # solver type, boundary type: the context's own
    info["maxwell_matrix_size"] = int(nxt())
    info["ext_j"] = float(nxt())
    for key in ("V0", "E0", "B0"):
        info[key] = [float(nxt()) for _ in range(3)]
    return info


def test_backup_files_match_reference(tmp_path):
    d = np.load(os.path.join(HERE, "golden", "backup_ref.npz"))
    w = wl.mini()
    print("case", w.name, "backup of the initial state", flush=True)
    st = wl.InitialState(d["in/species"].astype(np.int32), d["in/ids"].astype(np.int32), d["in/x"], d["in/p"],
                         d["in/E"], d["in/B"])
    sim = Simulation(w, precision="fp64")
    sim.load_state(st)
    ref = {k[5:]: d[k].tobytes().decode() for k in d.files if k.startswith("file/")}
    info = parse_general(ref["general"])
    info["species_type"] = [{"electron": 0, "proton": 1}[s.name] for s in w.species]
    info["particle_dx"] = w.dx / 2
    names = {k: str(tmp_path / (k + ".dat")) for k in ("general", "E", "B", "particles")}
    sim.outputSimulationBackup(names["general"], names["E"], names["B"], names["particles"], **info)
    for k, path in names.items():
        with open(path) as f:
            mine = f.read()
        if mine != ref[k]:
            a, b = mine.splitlines(), ref[k].splitlines()
            bad = next((i for i, (x, y) in enumerate(zip(a, b)) if x != y), min(len(a), len(b)))
            pytest.fail("%s differs at line %d: %r vs %r" % (k, bad, a[bad] if bad < len(a) else None,
                                                            b[bad] if bad < len(b) else None))
    sim.close()
