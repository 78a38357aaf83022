"""Benchmark of the implicit PIC step on one MI355X.

Metric: particle-steps/s including every deposit (one particle advanced one full
implicit step: moments, field solve, push, re-sort and post-push density), plus the
GMRES time per Maxwell solve to the 1e-8 true residual, learned preconditioner against
plain GMRES. Workload: the one-rank production box, 128^3 cells, 64 particles per cell
per species (electron-proton, m_p/m_e = 100, kT/mc^2 = 0.01, omega_ce/omega_pe = 0.5,
dt = 0.1/omega_pe), synthetic seeded initial state generated on the GPU.

Default run: the chosen form (FP32 mover/deposit, mixed-precision GMRES) is timed for `--steps`
steps from the initial state and gives `value`; the learned preconditioner is then trained on
the box's own solves (2000 pairs logged over the following steps, untimed), and the systems of
the next `--held-out-steps` steps are solved by plain and by learned-preconditioned GMRES; the
plain FP64 step (FP64 GMRES to the same residual) and the explicit step are timed beside it
from the same initial state; the CPU baseline is the FP64 oracle on the reference's CPU case.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 128] [--ppc 64]
                       [--precision fp32|fp64] [--solver mixed|fp64] [--learned off|compare|use]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

from picpp_amd._lib import PicError  # noqa: E402
from picpp_amd.simulation import open_synthetic_box  # noqa: E402


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--stop-rule", default="true_residual", choices=["reference", "true_residual"])
    ap.add_argument("--ortho", default="mgs", choices=["mgs", "cgs2"])
    ap.add_argument("--solver", default="mixed", choices=["fp64", "mixed"],
                    help="arithmetic of the Maxwell solve in the timed steps")
    ap.add_argument("--learned", default="compare", choices=["off", "compare", "use"],
                    help="learned preconditioner: off; trained and compared with plain GMRES on held-out systems "
                         "(compare); or also used in the timed steps (use)")
    ap.add_argument("--train-on", default="production", choices=["production", "small"],
                    help="learned preconditioner's training pairs: the first --train-pairs that the benchmark box's own "
                         "solves log (production), or a train_cells^3 box of the same physics (small)")
    ap.add_argument("--held-out-steps", type=int, default=3,
                    help="after the timed steps, the systems of this many further steps are solved by plain and by "
                         "learned-preconditioned GMRES (none seen in training)")
    ap.add_argument("--train-cells", type=int, default=32, help="cells per axis of the small training box")
    ap.add_argument("--train-pairs", type=int, default=2000)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--no-fp64-compare", action="store_true", help="skip timing the plain FP64 step beside the chosen form")
    ap.add_argument("--drift-steps", type=int, default=0,
                    help="also run the chosen form and the plain FP64 step this many steps from the same state and "
                         "report their total-energy difference")
    ap.add_argument("--weibel", type=int, default=0, metavar="STEPS",
                    help="only run the counterstreaming Weibel case (e-/e+ beams along +/-z, drift gamma 5) for STEPS steps "
                         "in the chosen form and in FP64, and compare magnetic-energy growth and total energy")
    ap.add_argument("--weibel-cells", type=int, default=192)
    ap.add_argument("--weibel-ppc", type=int, default=32, help="particles per cell per species per beam")
    ap.add_argument("--shock", type=int, default=0, metavar="STEPS",
                    help="run the relativistic shock set-up (reflecting wall, flow injected at the right end) for STEPS "
                         "steps and report its throughput and density pile-up instead of the default benchmark")
    ap.add_argument("--explicit", type=int, default=0, metavar="STEPS",
                    help="only time STEPS steps of the explicit (BUNEMAN / Tristan) FP64 step on the benchmark box "
                         "(--cells, --ppc) instead of the default benchmark")
    ap.add_argument("--explicit-steps", type=int, default=5,
                    help="with the default benchmark, also time this many explicit (BUNEMAN / Tristan) FP64 steps on the "
                         "same box and report them as explicit_step (0: skip)")
    ap.add_argument("--precond-dt-sweep", default=None, metavar="DT,DT,...",
                    help="only compare plain and learned-preconditioned mixed GMRES at each time step (in 1/omega_pe): "
                         "the network is trained at that dt on the train_cells^3 box, then both solve held-out systems of "
                         "the benchmark box (--cells, --ppc) to the 1e-8 true residual")
    ap.add_argument("--sweep-held-out", type=int, default=3, help="held-out systems per time step of --precond-dt-sweep")
    ap.add_argument("--precond-production", type=int, default=0, metavar="HELD_OUT",
                    help="only train the learned preconditioner on the benchmark box's own solves (--cells, --ppc; the "
                         "first --train-pairs pairs that the mixed-precision solver logs), then compare plain and learned "
                         "GMRES on the systems of the next HELD_OUT steps, each to the 1e-8 true residual")
    ap.add_argument("--shock-cells", type=int, default=1024, help="cells along x (64 x 64 across)")
    ap.add_argument("--shock-ppc", type=int, default=16, help="particles per cell per species at injection")
    ap.add_argument("--pressure", default=None, choices=["reference", "cell"],
                    help="pressure tensor deposit: the reference's (default, except --shock) or cell-local (default for "
                         "--shock: the reference's deposit gathers each x plane's pressure in the corner columns, which "
                         "drives the fields there without bound in the shock set-up)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB between two device buffers three times (the reference kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--no-held-out", action="store_true", help="skip the held-out solver comparison after the timed steps")
    ap.add_argument("--pmc-summary", default=None, metavar="DIR",
                    help="summarise FETCH_SIZE / WRITE_SIZE counter runs under DIR into profiles/pmc_traffic.json")
    ap.add_argument("--trace-summary", default=None, metavar="DIR",
                    help="per-kernel, per-grid-size dispatch statistics of the rocprofv3 kernel traces under DIR (markdown)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU baseline sample")
    return ap.parse_args()


def cpu_baseline(seconds: float):
    """The FP64 oracle (a single-threaded C++ restatement of the reference step) on a
    bounded sample of the same physics: the reference's CPU case, 32x32x8 cells with
    20 particles per cell per species, stepped until about `seconds` have passed."""
    from oracle.oracle import OracleSim
    from picpp_amd import workloads as wl

    w = wl.config1()
    st = wl.make_initial_state(w)
    o = OracleSim(w, st)
    o.initial_sequence()
    o.step(1)  # warm
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    npart = len(st.species)
    out = {"value": npart * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
           "sample": "oracle FP64 step (single-threaded), 32x32x8 cells, 20 ppc per species (%d particles), %d steps in %.1f s"
                     % (npart, steps, el)}
    # the compiled reference itself cannot run on this machine (its sources are not here); its
    # timing on the same workload, taken where it was built, is committed beside the profiles
    path = os.path.join(ROOT, "profiles", "reference_cpu_implicit_32x32x8_ppc20.json")
    if os.path.exists(path):
        with open(path) as fh:
            ref = json.load(fh)
        out["compiled_reference"] = {"cpu": ref["cpu"], "steps": ref["steps"],
                                     "particle_steps_per_s_by_ranks": {str(r["ranks"]): r["particle_steps_per_s"] for r in ref["runs"]},
                                     "source": "profiles/reference_cpu_implicit_32x32x8_ppc20.md"}
    return out


def train_network(args, torch, dev, w=None):
    """Trains the learned preconditioner on pairs logged by the mixed-precision solves of a
    train_cells^3 box with the same physics (by default the benchmark's, 16 ppc per species).
    The network is convolutional and periodic, so its parameters carry over to the production
    grid."""
    from picpp_amd import workloads as wl

    if w is None:
        w = wl.config2(n=(args.train_cells,) * 3, ppc=16, steps=0)
    sim = open_synthetic_box(w, torch, dev, precision="fp32", stop_rule="true_residual", solver="mixed")
    sim.precond_create()
    sim.precond_log(args.train_pairs)
    t0 = time.perf_counter()
    steps = 0
    while sim.precond_log(-1) < args.train_pairs and steps < 4000:
        sim.step()
        steps += 1
    t_log = time.perf_counter() - t0
    npairs = sim.precond_log(-1)
    _, identity = sim.precond_train(0, 8, 2e-3, 0.1)
    t0 = time.perf_counter()
    tl, hl = sim.precond_train(args.train_epochs, 8, 2e-3, 0.1)
    t_train = time.perf_counter() - t0
    params, mu_scale = sim.precond_params()
    sim.close()
    print("network: %d pairs from %d steps of a %d^3 box in %.1f s; %d epochs in %.1f s; held-out loss %.3e (identity %.3e)"
          % (npairs, steps, args.train_cells, t_log, args.train_epochs, t_train, hl, identity), file=sys.stderr, flush=True)
    return params, mu_scale, {"box": [args.train_cells] * 3, "pairs": npairs, "solves": steps, "log_seconds": t_log,
                              "epochs": args.train_epochs, "train_seconds": t_train, "train_loss": tl,
                              "holdout_loss": hl, "identity_holdout_loss": identity}


def compare_solvers(sim, torch, learned, reps=3):
    """Solves the current step's Maxwell system (held out from training) with plain mixed
    GMRES and, if available, with the learned preconditioner; each to the 1e-8 true residual."""
    sim.updateElectroMagneticParameters()
    sim.evaluateMaxwellEquationMatrix()
    rhs_ptr, n = sim.device_array("rhs")
    out = {}
    for name, use in (("plain", False), ("learned", True)):
        if use and not learned:
            continue
        if learned:
            sim.precond_use(use)
        runs = []
        for _ in range(reps):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            runs.append(sim.gmres_solve(rhs_ptr, x.data_ptr()))
        out[name] = {"ms": 1e3 * float(np.mean([r["seconds"] for r in runs])),
                     "iterations": float(np.mean([r["iterations"] for r in runs])),
                     "true_residual": float(max(r["true_residual"] for r in runs)),
                     "fallbacks": int(sum(r["fallbacks"] for r in runs))}
    if learned:
        sim.precond_use(False)
    return out


def precond_dt_sweep(args, torch, dev):
    """Plain against learned-preconditioned mixed GMRES as the implicit system stiffens: per time
    step, the network is trained at that dt (train_network on the train_cells^3 box), the
    benchmark box is stepped twice at that dt, and each of sweep_held_out systems (one per
    further step, none seen in training) is solved by both, to the 1e-8 true residual."""
    import dataclasses

    from picpp_amd import workloads as wl

    rows = []
    for dt in [float(v) for v in args.precond_dt_sweep.split(",")]:
        tw = dataclasses.replace(wl.config2(n=(args.train_cells,) * 3, ppc=16, steps=0), dt=dt)
        params, mu_scale, training = train_network(args, torch, dev, tw)
        w = dataclasses.replace(wl.config2(n=(args.cells,) * 3, ppc=args.ppc, steps=0), dt=dt)
        sim = open_synthetic_box(w, torch, dev, precision="fp32", stop_rule="true_residual", solver="mixed")
        sim.precond_create()
        sim.set_precond_params(params, mu_scale)
        sim.precond_use(False)
        held = []
        try:
            for _ in range(2):
                sim.step()
            for _ in range(args.sweep_held_out):
                held.append(compare_solvers(sim, torch, True))
                sim.step()
        except PicError as e:  # the step's own checks refuse this time step: reported, the sweep goes on
            print("dt %.3g: %s" % (dt, e), file=sys.stderr, flush=True)
            rows.append({"dt": dt, "training": training, "failed": str(e)})
            continue
        finally:
            sim.close()
            del sim
            torch.cuda.empty_cache()
        row = {"dt": dt, "training": training}
        for name in ("plain", "learned"):
            row[name] = {k: float(np.mean([h[name][k] for h in held])) for k in ("ms", "iterations", "true_residual")}
            row[name]["fallbacks"] = int(sum(h[name]["fallbacks"] for h in held))
        row["learned_over_plain_time"] = row["learned"]["ms"] / row["plain"]["ms"]
        print("dt %.3g: plain %.1f iterations %.3f ms, learned %.1f iterations %.3f ms" % (
            dt, row["plain"]["iterations"], row["plain"]["ms"], row["learned"]["iterations"], row["learned"]["ms"]),
              file=sys.stderr, flush=True)
        rows.append(row)
    return {"cells": [args.cells] * 3, "ppc_per_species": args.ppc, "held_out_per_dt": args.sweep_held_out, "rows": rows}


def log_and_train(sim, args, torch):
    """Trains the learned preconditioner of `sim` on its own solves: the box is stepped in the
    chosen form while the mixed-precision solver logs its first train_pairs pairs (every Arnoldi
    step gives (A v_j, v_j), every solve (b, x)); the network is trained on them with 10 % held
    out, and the logged pairs are freed."""
    sim.precond_create()
    sim.precond_log(args.train_pairs)
    t0 = time.perf_counter()
    steps = 0
    while sim.precond_log(-1) < args.train_pairs and steps < 4000:
        sim.step()
        steps += 1
    torch.cuda.synchronize()
    t_log = time.perf_counter() - t0
    npairs = sim.precond_log(-1)
    print("logged %d pairs from %d steps in %.1f s" % (npairs, steps, t_log), file=sys.stderr, flush=True)
    _, identity = sim.precond_train(0, 8, 2e-3, 0.1)
    t0 = time.perf_counter()
    tl, hl = sim.precond_train(args.train_epochs, 8, 2e-3, 0.1)
    t_train = time.perf_counter() - t0
    print("trained %d epochs in %.1f s: held-out loss %.3e (identity %.3e)" % (args.train_epochs, t_train, hl, identity),
          file=sys.stderr, flush=True)
    sim.precond_log(0)  # frees the logged pairs
    return {"box": list(sim.w.n), "pairs": npairs, "solves": steps, "log_seconds": t_log, "epochs": args.train_epochs,
            "train_seconds": t_train, "train_loss": tl, "holdout_loss": hl, "identity_holdout_loss": identity}


def held_out_steps(sim, torch, learned, count):
    """The systems of the next `count` steps, none seen in training, each solved by plain and (if
    trained) learned-preconditioned mixed GMRES to the 1e-8 true residual: per step and mean."""
    held = []
    for k in range(count):
        h = compare_solvers(sim, torch, learned)
        held.append(h)
        print("held-out step %d: plain %.1f iterations %.3f ms" % (k + 1, h["plain"]["iterations"], h["plain"]["ms"])
              + (", learned %.1f iterations %.3f ms (fallbacks %d)" % (h["learned"]["iterations"], h["learned"]["ms"],
                                                                       h["learned"]["fallbacks"]) if learned else ""),
              file=sys.stderr, flush=True)
        if k + 1 < count:
            sim.step()
    summary = {}
    for name in ("plain", "learned"):
        if name not in held[0]:
            continue
        summary[name] = {q: float(np.mean([h[name][q] for h in held])) for q in ("ms", "iterations")}
        summary[name]["true_residual_max"] = float(max(h[name]["true_residual"] for h in held))
        summary[name]["fallbacks"] = int(sum(h[name]["fallbacks"] for h in held))
    out = {"steps": held, "mean": summary}
    if "learned" in summary:
        out["learned_over_plain_time"] = summary["learned"]["ms"] / summary["plain"]["ms"]
    return out


def precond_production(args, torch, dev):
    """The learned preconditioner trained on the production box itself (log_and_train on the
    benchmark box, --cells, --ppc), then plain against learned-preconditioned GMRES on the
    systems of the next `precond_production` steps, each to the 1e-8 true residual."""
    from picpp_amd import workloads as wl

    w = wl.config2(n=(args.cells,) * 3, ppc=args.ppc, steps=0)
    sim = open_synthetic_box(w, torch, dev, precision="fp32", stop_rule="true_residual", solver="mixed")
    npart = sim.num_particles()
    training = log_and_train(sim, args, torch)
    held = held_out_steps(sim, torch, True, args.precond_production)
    sim.close()
    return {"cells": [args.cells] * 3, "ppc_per_species": args.ppc, "particles": npart, "dt": w.dt, "training": training,
            "held_out": held}


def time_steps(sim, args):
    for i in range(args.warmup):
        sim.step()
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    import torch
    sim._chk(sim.lib.pic_profile_enable(sim.h, 1), "pic_profile_enable")
    solves = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        solves.append(sim.step())
    torch.cuda.synchronize()
    return time.perf_counter() - t0, solves


def read_profile(sim, steps):
    _D = ctypes.POINTER(ctypes.c_double)
    names = ctypes.create_string_buffer(64 * 32)
    ms = np.zeros(64)
    cn = np.zeros(64, np.int64)
    by = np.zeros(64)
    k = sim.lib.pic_profile_read(sim.h, 64, names, ms.ctypes.data_as(_D),
                                 cn.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), by.ctypes.data_as(_D))
    sim._chk(k, "pic_profile_read")
    kernels = {}
    for i in range(k):
        nm = names.raw[32 * i:32 * i + 32].split(b"\0")[0].decode()
        kernels[nm] = {"ms_total": ms[i], "launches": int(cn[i]), "bytes_total": by[i]}
    for nm, kv in sorted(kernels.items(), key=lambda q: -q[1]["ms_total"]):
        bw = kv["bytes_total"] / (kv["ms_total"] * 1e-3) / 1e9 if kv["ms_total"] > 0 else 0
        print("kernel %-18s %8.3f ms/step  %5d launches  %7.1f GB/s algorithmic" %
              (nm, kv["ms_total"] / steps, kv["launches"], bw), file=sys.stderr)
    return kernels


def step_roofline(npart, n, ms, args, state_bytes=28):
    """The whole step against the HBM roofline, per the step's algorithmic bytes: each particle's
    state read once and written once (2 S: 56 B in FP32, 96 B in FP64) plus 45 FP64 words per node
    (E^{n+theta}, B^n, E^{n+1}, B^{n+1} read; J, mu, rhohat, Pi, rho and the two species' five
    density planes written). With the committed counter summary of this workload, also the
    measured HBM bytes of one step and their ratio to the algorithmic bytes."""
    nodes = n[0] * n[1] * n[2]
    alg = npart * 2.0 * state_bytes + nodes * 360.0
    out = {"algorithmic_bytes": alg, "achieved_gbs": alg / (ms * 1e-3) / 1e9,
           "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "unit": "GB/s", "peak": HBM_PEAK_GBS}
    rec = pmc_record(args) if state_bytes == 28 else None
    if rec and "step_traffic_bytes" in rec:
        out["counter_traffic_bytes"] = rec["step_traffic_bytes"]
        out["counter_traffic_ratio"] = rec["step_traffic_bytes"] / alg
    return out


def energy_drift(args, w, torch, dev):
    """Total energy of the chosen form and of the plain FP64 step over drift_steps steps from
    the same initial state: the largest relative difference between the two series, and each
    series' own largest relative change from its initial value."""
    series = {}
    for name, kw in (("chosen", dict(precision=args.precision, stop_rule=args.stop_rule, solver=args.solver)),
                     ("fp64", dict(precision="fp64", stop_rule="true_residual", solver="fp64"))):
        sim = open_synthetic_box(w, torch, dev, pressure=args.pressure or "reference", **kw)
        e = [sim.updateEnergy()[3]]
        for i in range(args.drift_steps):
            sim.step()
            e.append(sim.updateEnergy()[3])
            if (i + 1) % 50 == 0:
                print("drift %s: step %d" % (name, i + 1), file=sys.stderr, flush=True)
        sim.close()
        del sim
        torch.cuda.empty_cache()
        series[name] = np.array(e)
    a, b = series["chosen"], series["fp64"]
    return {"steps": args.drift_steps, "max_rel_difference": float(np.max(np.abs(a - b) / np.abs(b))),
            "chosen_rel_change": float(np.max(np.abs(a - a[0]) / abs(a[0]))),
            "fp64_rel_change": float(np.max(np.abs(b - b[0]) / abs(b[0])))}


PMC_PATH = os.path.join(ROOT, "profiles", "pmc_traffic.json")
# the step kernels' names in the counter runs, by the bench's per-kernel names
PMC_SLOTS = {"push": ("k_push32<",), "density_cell_moments": ("k_cell_moments32<true>",),
             "node_moments": ("k_node_moments32",), "permute": ("k_permute32",),
             "cell_moments": ("k_cell_moments32<false>",), "density": ("k_density32",)}


def pmc_record(args):
    """The committed counter summary, when it was taken on this workload (FP32, 128^3, 64 ppc)."""
    if args.precision != "fp32" or args.cells != 128 or args.ppc != 64 or not os.path.exists(PMC_PATH):
        return None
    with open(PMC_PATH) as fh:
        return json.load(fh)


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE counter passes
    (scaled by a plain copy kernel's counters), when they were taken on this workload."""
    rec = pmc_record(args)
    rec = rec["kernels"].get(kernel) if rec else None
    return rec["traffic_bytes"] if rec else None


def _counter_rows(path):
    """(dispatch id, kernel name, counter value) of every row of the rocprofv3 counter CSVs under path."""
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r.get("Dispatch_Id", r.get("Correlation_Id", 0))), r["Kernel_Name"],
                             float(r["Counter_Value"])))
    if not rows:
        raise SystemExit("no counter CSV under %s" % path)
    return rows


def _per_dispatch(path):
    tot = {}
    for d, name, v in _counter_rows(path):
        k = (d, name)
        tot[k] = tot.get(k, 0.0) + v  # a counter may come per XCD / instance
    return tot


def short_kernel_name(nm):
    nm = nm.replace("(anonymous namespace)::", "")
    depth, out = 0, []
    for ch in nm:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0 and out:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).replace("void ", "", 1)


def trace_summary(d, top=40, skip=2, first=5):
    """Dispatch statistics per kernel and grid size from the rocprofv3 kernel-trace CSVs under d:
    the launches of one kernel on boxes of different sizes (the bench's 128^3 box, the learned
    preconditioner's 32^3 box) are kept apart. For the launches of 16 M work-items and more
    (the production box's particle and grid kernels) a second table gives the average of
    launches skip + 1 .. skip + first of each kernel: in the default run those are the chosen
    form's timed steps (after 2 warmup steps; the FP32 kernels' first launches), which come
    before the steps that log the learned preconditioner's training pairs."""
    import csv
    import glob

    stats = {}
    launches = {}
    for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                grid = 1
                for ax in ("X", "Y", "Z"):
                    grid *= int(r.get("Grid_Size_" + ax, 1) or 1)
                if "Grid_Size" in r:
                    grid = int(r["Grid_Size"])
                k = (short_kernel_name(r["Kernel_Name"]), grid)
                ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                c, t = stats.get(k, (0, 0))
                stats[k] = (c + 1, t + ns)
                if grid >= (1 << 24):
                    launches.setdefault(k, []).append((int(r["Start_Timestamp"]), ns))
    total = sum(t for c, t in stats.values())
    rows = sorted(stats.items(), key=lambda kv: -kv[1][1])
    lines = ["| Kernel | Grid (work-items) | Calls | Average ms | Total ms | % |", "|---|---|---|---|---|---|"]
    for (nm, grid), (c, t) in rows[:top]:
        lines.append("| `%s` | %d | %d | %.3f | %.3f | %.1f |" % (nm[:80], grid, c, t / c / 1e6, t / 1e6, 100.0 * t / total))
    big = [kv for kv in rows[top:] if kv[0][1] >= (1 << 24)]
    if big:  # the production-size launches (16 M work-items and more) not listed above
        lines.append("| *(launches of 16 M work-items and more, below the top %d)* | | | | | |" % top)
        for (nm, grid), (c, t) in big:
            lines.append("| `%s` | %d | %d | %.3f | %.3f | %.1f |" % (nm[:80], grid, c, t / c / 1e6, t / 1e6, 100.0 * t / total))
    lines.append("")
    lines.append("All dispatches: %.1f ms, %d kernel and grid-size pairs." % (total / 1e6, len(stats)))
    lines.append("")
    lines.append("Launches %d to %d of each kernel of 16 M work-items and more (by start time):" % (skip + 1, skip + first))
    lines.append("")
    lines.append("| Kernel | Grid (work-items) | Launches averaged | Average ms |")
    lines.append("|---|---|---|---|")
    window = {k: sorted(v)[skip:skip + first] or sorted(v)[:first] for k, v in launches.items()}
    for (nm, grid), ls in sorted(window.items(), key=lambda kv: -sum(x[1] for x in kv[1])):
        lines.append("| `%s` | %d | %d | %.3f |" % (nm[:80], grid, len(ls), sum(x[1] for x in ls) / len(ls) / 1e6))
    return "\n".join(lines)


def pmc_summary(d, steps_a, steps_b):
    """Per-launch and per-step HBM traffic from six counter runs under d: FETCH_SIZE and WRITE_SIZE
    of the timed step at steps_a and at steps_b steps (the difference divided by steps_b - steps_a is
    one step's traffic, since everything else in the two runs is the same), and of a plain 1 GiB
    device copy (the scale of each counter)."""
    cf = _per_dispatch(os.path.join(d, "copy_fetch"))
    cw = _per_dispatch(os.path.join(d, "copy_write"))
    gib = float(1 << 30)
    copy_f = [v for (i, nm), v in cf.items() if "copyBuffer" in nm or "copy" in nm.lower()]
    copy_w = [v for (i, nm), v in cw.items() if "copyBuffer" in nm or "copy" in nm.lower()]
    copy_f = float(np.median(sorted(copy_f)[-3:]))
    copy_w = float(np.median(sorted(copy_w)[-3:]))
    ff = gib / (copy_f * 1024.0)  # bytes per counted byte, each counter in KiB
    fw = gib / (copy_w * 1024.0)
    runs = {}
    for tag, st in (("a", steps_a), ("b", steps_b)):
        f = _per_dispatch(os.path.join(d, "fetch_%d" % st))
        w = _per_dispatch(os.path.join(d, "write_%d" % st))
        runs[tag] = (f, w)

    def total(f, w):
        return sum(v * 1024.0 * ff for v in f.values()) + sum(v * 1024.0 * fw for v in w.values())

    step_bytes = (total(*runs["b"]) - total(*runs["a"])) / (steps_b - steps_a)

    def by_name(f, w):
        out = {}
        for (i, nm), v in f.items():
            out[nm] = out.get(nm, 0.0) + v * 1024.0 * ff
        for (i, nm), v in w.items():
            out[nm] = out.get(nm, 0.0) + v * 1024.0 * fw
        return out

    na, nb_ = by_name(*runs["a"]), by_name(*runs["b"])
    per_kernel = {nm: (nb_.get(nm, 0.0) - na.get(nm, 0.0)) / (steps_b - steps_a) for nm in set(na) | set(nb_)}
    per_kernel = dict(sorted(((k[:90], v) for k, v in per_kernel.items() if abs(v) > 1e7), key=lambda kv: -kv[1]))
    f, w = runs["b"]
    kernels = {}
    for slot, pats in PMC_SLOTS.items():
        ids = sorted({k for k in f if any(p in k[1] for p in pats)})
        if not ids:
            continue
        fk = sum(f[k] for k in ids)
        wk = sum(w.get(k, 0.0) for k in ids)
        kernels[slot] = {"fetch_kib_raw": fk / len(ids), "write_kib_raw": wk / len(ids),
                         "traffic_bytes": (fk * ff + wk * fw) * 1024.0 / len(ids), "dispatches": len(ids)}
    return {"fetch_factor": ff, "write_factor": fw,
            "calibration": "1 GiB device-to-device copy: FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB per dispatch" % (copy_f, copy_w),
            "step_traffic_bytes": step_bytes, "steps": [steps_a, steps_b], "kernels": kernels,
            "step_traffic_by_kernel": per_kernel}


def weibel_check(args, torch, dev):
    """Counterstreaming Weibel filamentation: the chosen form (FP32 mover/deposit, mixed GMRES
    with the learned preconditioner unless --learned off, trained on a train_cells^3 box of the
    same beams) and the plain FP64 step from the same state. The magnetic energy's growth rate
    is the steepest least-squares slope of ln W_B over 20-step windows; total energy is compared
    step by step."""
    from picpp_amd import workloads as wl

    w = wl.weibel(n=(args.weibel_cells,) * 3, ppc=args.weibel_ppc, steps=args.weibel)
    learned = args.learned != "off" and args.solver == "mixed" and args.precision == "fp32"
    training = None
    if learned:
        params, mu_scale, training = train_network(args, torch, dev, wl.weibel(n=(args.train_cells,) * 3, ppc=8, steps=0))
    res = {}
    solver_stats = {}
    for name, kw in (("chosen", dict(precision=args.precision, stop_rule=args.stop_rule, solver=args.solver)),
                     ("fp64", dict(precision="fp64", stop_rule="true_residual", solver="fp64"))):
        sim = open_synthetic_box(w, torch, dev, pressure=args.pressure or "reference", **kw)
        if name == "chosen" and learned:
            sim.precond_create()
            sim.set_precond_params(params, mu_scale)
            sim.precond_use(True)
        npart = sim.num_particles()
        e = [sim.updateEnergy()[:4]]
        its, secs, falls = [], [], 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.weibel):
            info = sim.step()
            its.append(info["iterations"])
            secs.append(info["seconds"])
            falls += info["fallbacks"]
            e.append(sim.updateEnergy()[:4])
            if (i + 1) % 25 == 0:
                print("weibel %s: step %d" % (name, i + 1), file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        sim.close()
        del sim
        torch.cuda.empty_cache()
        res[name] = (np.array(e), el, npart)
        solver_stats[name] = {"gmres_iterations_per_solve": float(np.mean(its)), "gmres_ms_per_solve": 1e3 * float(np.mean(secs)),
                              "fallbacks": int(falls)}
    t = np.arange(args.weibel + 1) * w.dt
    out = {"cells": [args.weibel_cells] * 3, "particles": res["chosen"][2], "steps": args.weibel, "dt": w.dt,
           "pressure": args.pressure or "reference", "learned_preconditioner": learned, "training": training}
    for name, (e, el, npart) in res.items():
        lb = np.log(np.maximum(e[:, 2], 1e-300))
        # B starts at zero and the first steps only bring up the particle noise level: the
        # search for the exponential phase starts after the first 10 % of the steps (short runs
        # fit shorter windows)
        k0 = max(1, len(t) // 10)
        win = max(2, min(20, len(t) - k0 - 1))
        slopes = [np.polyfit(t[k:k + win], lb[k:k + win], 1)[0] for k in range(k0, len(t) - win + 1)]
        k = k0 + int(np.argmax(slopes))
        out[name] = {**solver_stats[name], "ms_per_step": 1e3 * el / args.weibel, "particle_steps_per_s": npart * args.weibel / el,
                     "magnetic_energy_growth_rate": float(slopes[k - k0]), "growth_window_steps": [k, k + win],
                     "magnetic_energy_every_10_steps": [float(v) for v in e[::10, 2]],
                     "magnetic_energy_final": float(e[-1, 2]), "total_energy_rel_change": float(np.max(np.abs(e[:, 3] - e[0, 3]) / e[0, 3]))}
    a, b = res["chosen"][0], res["fp64"][0]
    out["growth_rate_rel_difference"] = abs(out["chosen"]["magnetic_energy_growth_rate"] / out["fp64"]["magnetic_energy_growth_rate"] - 1)
    out["total_energy_max_rel_difference"] = float(np.max(np.abs(a[:, 3] - b[:, 3]) / np.abs(b[:, 3])))
    out["magnetic_energy_max_rel_difference"] = float(np.max(np.abs(a[:, 2] - b[:, 2]) / np.maximum(np.abs(b[:, 2]), 1e-300)))
    return out


def explicit_run(args, torch, dev):
    """The reference's BUNEMAN step (Yee fields, Tristan mover, zig-zag current) in FP64 on the
    benchmark's plasma: particle-steps/s and the per-kernel times of the timed steps."""
    from picpp_amd import workloads as wl

    w = wl.config2(n=(args.cells,) * 3, ppc=args.ppc, steps=args.explicit)
    sim = open_synthetic_box(w, torch, dev, precision="fp64", scheme="explicit")
    sim.load_yee(wl.yee_initial_fields(w))
    npart = sim.num_particles()
    print("explicit: %s, %d particles" % (w.name, npart), file=sys.stderr, flush=True)
    for i in range(args.warmup):
        sim.explicit_step()
    sim._chk(sim.lib.pic_profile_enable(sim.h, 1), "pic_profile_enable")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.explicit):
        sim.explicit_step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kernels = read_profile(sim, args.explicit)
    sim.close()
    return {"cells": [args.cells] * 3, "particles": npart, "steps": args.explicit, "precision": "fp64",
            "ms_per_step": 1e3 * el / args.explicit, "particle_steps_per_s": npart * args.explicit / el,
            "kernels_ms_per_step": {nm: kv["ms_total"] / args.explicit for nm, kv in kernels.items()}}


def shock_run(args, torch, dev):
    """The relativistic shock set-up: a cold e-/p+ flow (gamma0 10, mass ratio 64,
    kT/mc^2 = 1e-4, sigma 0.1) onto the reflecting wall, injected at the right end every step.
    Reports the step time, the particle count and the electron density profile along x
    (averaged over y, z), whose pile-up behind the wall makes per-cell counts strongly
    nonuniform for the sorted deposits."""
    from picpp_amd import workloads as wl

    w = wl.shock(n=(args.shock_cells, 64, 64), ppc=args.shock_ppc, steps=args.shock)
    pressure = args.pressure or "cell"
    sim = open_synthetic_box(w, torch, dev, precision=args.precision, stop_rule=args.stop_rule, solver=args.solver, pressure=pressure)
    n0 = sim.num_particles()
    print("shock: %s, %d particles, precision %s, solver %s" % (w.name, n0, args.precision, args.solver), file=sys.stderr,
          flush=True)
    iters, secs, counts, energies = [], [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    paused = 0.0
    prof_steps = min(20, args.shock)
    for i in range(args.shock):
        if i == args.shock - prof_steps:  # per-kernel event timing over the last steps
            sim._chk(sim.lib.pic_profile_enable(sim.h, 1), "pic_profile_enable")
        s = sim.step()
        iters.append(s["iterations"])
        secs.append(s["seconds"])
        if (i + 1) % 25 == 0:  # bookkeeping outside the timed work
            t1 = time.perf_counter()
            counts.append(sim.num_particles())
            energies.append([float(v) for v in sim.updateEnergy()[:4]])
            paused += time.perf_counter() - t1
            print("shock: step %d, %d particles, energy (particles, E, B) %.4g %.4g %.4g, gmres %d, %.1f ms per step so far" % (
                i + 1, counts[-1], energies[-1][0], energies[-1][1], energies[-1][2], s["iterations"],
                1e3 * (time.perf_counter() - t0 - paused) / (i + 1)), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0 - paused
    n1 = sim.num_particles()
    kernels = read_profile(sim, prof_steps)
    dens = sim.array("density")
    sim.close()
    del sim
    torch.cuda.empty_cache()
    ne = dens[2:w.n[0] + 2, 2:w.n[1] + 2, 2:w.n[2] + 2, 0].mean(axis=(1, 2))  # electron concentration per x cell
    upstream = float(np.median(ne[len(ne) // 2:]))
    return {"cells": list(w.n), "ppc_per_species": args.shock_ppc, "gamma0": 10.0, "steps": args.shock, "dt": w.dt,
            "precision": args.precision, "solver": args.solver, "pressure": pressure, "particles_initial": n0,
            "particles_final": n1,
            "particle_count_every_25_steps": counts, "energy_every_25_steps": energies, "ms_per_step": 1e3 * el / args.shock,
            "particle_steps_per_s": 0.5 * (n0 + n1) * args.shock / el,
            "gmres_iterations_per_solve": float(np.mean(iters)), "gmres_ms_per_solve": 1e3 * float(np.mean(secs)),
            "electron_density_upstream": upstream, "electron_density_max_over_upstream": float(ne.max() / upstream),
            "electron_density_profile_every_8_cells": [float(v) for v in ne[::8]],
            "kernels_ms_per_step_last_%d_steps" % prof_steps: {nm: kv["ms_total"] / prof_steps for nm, kv in kernels.items()}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    from picpp_amd import workloads as wl

    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    if args.pmc_summary:
        rec = pmc_summary(args.pmc_summary, 1, 3)
        with open(PMC_PATH, "w") as fh:
            json.dump(rec, fh, indent=1)
        print(json.dumps(rec, indent=1))
        return
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copied 1 GiB three times", file=sys.stderr)
        return
    if args.weibel > 0:
        print(json.dumps({"weibel": weibel_check(args, torch, dev)}))
        return
    if args.shock > 0:
        print(json.dumps({"shock": shock_run(args, torch, dev)}))
        return
    if args.precond_dt_sweep:
        print(json.dumps({"precond_dt_sweep": precond_dt_sweep(args, torch, dev)}))
        return
    if args.precond_production > 0:
        print(json.dumps({"precond_production": precond_production(args, torch, dev)}))
        return
    if args.explicit > 0:
        print(json.dumps({"explicit": explicit_run(args, torch, dev)}))
        return
    learned = args.learned != "off" and args.solver == "mixed" and args.precision == "fp32"
    training = None
    if learned and args.train_on == "small":
        params, mu_scale, training = train_network(args, torch, dev)
    n = (args.cells,) * 3
    w = wl.config2(n=n, ppc=args.ppc, steps=args.steps)
    sim = open_synthetic_box(w, torch, dev, precision=args.precision, stop_rule=args.stop_rule, ortho=args.ortho, solver=args.solver,
                   pressure=args.pressure or "reference")
    npart = sim.num_particles()
    print("bench: %s, %d particles, precision %s, solver %s" % (w.name, npart, args.precision, args.solver), file=sys.stderr,
          flush=True)
    def setup_learned():
        if args.train_on == "production":  # the box's own solves (untimed)
            t = log_and_train(sim, args, torch)
        else:
            t = training
            sim.precond_create()
            sim.set_precond_params(params, mu_scale)
        sim.precond_use(args.learned == "use")
        return t

    # With the network in the timed steps it is trained first; for the comparison alone the
    # timed steps start from the initial state, as the plain FP64 step's do, and the network
    # is trained on the solves of the steps that follow them.
    if learned and args.learned == "use":
        training = setup_learned()
    el, solves = time_steps(sim, args)
    for i, s in enumerate(solves):
        print("step %d: gmres %d iterations, %.3f ms" % (i, s["iterations"], 1e3 * s["seconds"]), file=sys.stderr)
    kernels = read_profile(sim, args.steps)
    dom = max(kernels, key=lambda q: kernels[q]["ms_total"])
    kd = kernels[dom]
    avg_ms = kd["ms_total"] / kd["launches"]
    per_launch = kd["bytes_total"] / kd["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom, args)
    energy = sim.updateEnergy()
    if learned and args.learned != "use":
        sim._chk(sim.lib.pic_profile_enable(sim.h, 0), "pic_profile_enable")
        training = setup_learned()
    held_out = (held_out_steps(sim, torch, learned, args.held_out_steps)
                if args.solver == "mixed" and not args.no_held_out and args.held_out_steps > 0 else None)
    sim.close()
    del sim
    torch.cuda.empty_cache()
    out = {
        "metric": "particle-steps/s incl. deposit; GMRES ms to 1e-8 residual (learned vs plain)",
        "value": npart * args.steps / el,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": "implicit e-/p+ step, %dx%dx%d cells, %d ppc per species (%d particles), %s mover/deposit, "
                               "%s GMRES%s, %s stop rule" % (
                                   n[0], n[1], n[2], args.ppc, npart, args.precision, args.solver,
                                   " + learned preconditioner" if args.learned == "use" and learned else "", args.stop_rule),
                   "cells": list(n), "particles": npart, "precision": args.precision, "solver": args.solver,
                   "learned_preconditioner_in_timed_steps": bool(learned and args.learned == "use")},
        "gmres": {"ms_per_solve": 1e3 * float(np.mean([s["seconds"] for s in solves])),
                  "iterations_per_solve": float(np.mean([s["iterations"] for s in solves])),
                  "true_residual": float(max(s["true_residual"] for s in solves)), "stop_rule": args.stop_rule,
                  "held_out": held_out, "training": training},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": per_launch},
        "step_roofline": step_roofline(npart, n, 1e3 * el / args.steps, args),
        "kernels_ms_per_step": {nm: kv["ms_total"] / args.steps for nm, kv in kernels.items()},
        "energy": {"particles": energy[0], "E": energy[1], "B": energy[2], "total": energy[3]},
    }
    if not args.no_fp64_compare and args.precision == "fp32":
        sim = open_synthetic_box(w, torch, dev, precision="fp64", stop_rule="true_residual", solver="fp64",
                       pressure=args.pressure or "reference")
        el64, solves64 = time_steps(sim, args)
        sim.close()
        del sim
        torch.cuda.empty_cache()
        out["fp64_step"] = {"ms_per_step": 1e3 * el64 / args.steps, "value": npart * args.steps / el64,
                            "gmres_ms_per_solve": 1e3 * float(np.mean([s["seconds"] for s in solves64])),
                            "stop_rule": "true_residual", "speedup_of_chosen_form": el64 / el,
                            "step_roofline": step_roofline(npart, n, 1e3 * el64 / args.steps, args, state_bytes=48)}
        print("fp64 step: %.2f ms (chosen form %.2f ms, %.2fx)" % (1e3 * el64 / args.steps, 1e3 * el / args.steps, el64 / el),
              file=sys.stderr)
    if args.explicit_steps > 0 and args.precision == "fp32":
        # the reference's default (BUNEMAN) step on the same box, FP64, reported beside the headline
        ex = argparse.Namespace(**vars(args))
        ex.explicit, ex.warmup = args.explicit_steps, 1
        out["explicit_step"] = explicit_run(ex, torch, dev)
        torch.cuda.empty_cache()
        print("explicit step: %.2f ms" % out["explicit_step"]["ms_per_step"], file=sys.stderr, flush=True)
    if args.drift_steps > 0:
        out["energy_drift"] = energy_drift(args, w, torch, dev)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
