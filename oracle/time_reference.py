"""Times the compiled reference's implicit step on this machine's CPU cores (oracle/_ref/ref_timing,
built by oracle/Makefile.ref from the reference sources) on the reference's CPU case: 32x32x8
cells, 20 particles per cell per species (327,680 particles), the state of picpp_amd.workloads.config1
(seed 42), at the given MPI rank counts (a slab decomposition along x), and prints a markdown
summary with the per-phase times (and, with --out F.md, the same figures in F.json, which bench.py
reports beside its CPU baseline).

Measurement infrastructure only; it runs where the reference sources were built (the GPU machines
never have them).

Usage: python oracle/time_reference.py [--steps 50] [--ranks 1 8] [--out profiles/FILE.md]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import make_golden  # noqa: E402  (the harness input writer)
from picpp_amd import workloads as wl  # noqa: E402

TIMER = os.path.join(HERE, "_ref", "ref_timing")
MPIEXEC = "/opt/conda/bin/mpiexec"


def cpu_name():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run(ranks, fin, steps, td):
    cmd = [MPIEXEC, "-n", str(ranks), TIMER, fin, str(steps)]
    t0 = time.perf_counter()
    res = subprocess.run(cmd, check=True, capture_output=True, text=True, cwd=td)
    wall = time.perf_counter() - t0
    rec = {"ranks": ranks, "wall_s": wall, "phases": {}, "command": " ".join(cmd[:3] + ["oracle/_ref/ref_timing", "in.bin", str(steps)])}
    for line in res.stdout.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "phase":
            rec["phases"][f[1]] = float(f[2])
        elif f[0] in ("loop", "pre_loop", "particle_steps_per_s"):
            rec[f[0]] = float(f[1])
        elif f[0] == "particles_after":
            rec["particles_after"] = int(f[1])
            rec["energy"] = float(f[3])
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--ranks", type=int, nargs="+", default=[1, 8])
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    if not os.path.exists(TIMER):
        subprocess.run(["make", "-C", HERE, "-f", "Makefile.ref", "-j8"], check=True)
    w = wl.config1()
    st = wl.make_initial_state(w)
    recs = []
    with tempfile.TemporaryDirectory() as td:
        for d in make_golden.REF_DIRS:
            os.makedirs(os.path.join(td, d), exist_ok=True)
        fin = os.path.join(td, "in.bin")
        make_golden.write_input(fin, w, st, args.steps)
        for r in args.ranks:
            rec = run(r, fin, args.steps, td)
            print("ranks %d: loop %.2f s, %.3e particle-steps/s" % (r, rec["loop"], rec["particle_steps_per_s"]),
                  file=sys.stderr, flush=True)
            recs.append(rec)
    npart = len(st.species)
    phases = list(recs[0]["phases"])
    lines = ["# Compiled reference, implicit step on the CPU: %dx%dx%d cells, %d ppc per species (%d particles), %d steps"
             % (w.n[0], w.n[1], w.n[2], w.species[0].ppc, npart, args.steps), "",
             "Machine: %s, %d logical CPUs. The reference's 3-D code is pure MPI (MPICH, one process per rank, "
             "no threads); the ranks form a slab decomposition along x." % (cpu_name(), os.cpu_count()), "",
             "Command (from the repository root, after `make -C oracle -f Makefile.ref`):", "", "```",
             "python oracle/time_reference.py --steps %d --ranks %s" % (args.steps, " ".join(map(str, args.ranks))),
             "```", "", "which runs, per rank count, `%s` on the input file of the workload's initial state "
             "(oracle/ref_timing.cpp; every phase ends with a barrier, so each time is the slowest rank's)." % recs[0]["command"],
             "", "| Phase | " + " | ".join("%d rank%s: s / step" % (r["ranks"], "" if r["ranks"] == 1 else "s") for r in recs) + " |",
             "|---|" + "---|" * len(recs)]
    for ph in phases:
        lines.append("| %s | " % ph + " | ".join("%.4f" % (r["phases"][ph] / args.steps) for r in recs) + " |")
    lines.append("| **step (loop wall-clock / steps)** | " + " | ".join("**%.4f**" % (r["loop"] / args.steps) for r in recs) + " |")
    lines.append("| particle-steps/s | " + " | ".join("%.3e" % r["particle_steps_per_s"] for r in recs) + " |")
    lines.append("| pre-loop sequence, s | " + " | ".join("%.3f" % r["pre_loop"] for r in recs) + " |")
    lines.append("| total energy after the last step | " + " | ".join("%.12e" % r["energy"] for r in recs) + " |")
    text = "\n".join(lines) + "\n"
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text)
        summary = {"workload": "implicit e-/p+ step, %dx%dx%d cells, %d ppc per species (%d particles)"
                               % (w.n[0], w.n[1], w.n[2], w.species[0].ppc, npart),
                   "steps": args.steps, "cpu": cpu_name(), "logical_cpus": os.cpu_count(),
                   "runs": [{"ranks": r["ranks"], "particle_steps_per_s": r["particle_steps_per_s"],
                             "s_per_step": r["loop"] / args.steps,
                             "phases_s_per_step": {k: v / args.steps for k, v in r["phases"].items()}} for r in recs]}
        with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
            json.dump(summary, fh, indent=1)
    print(text)


if __name__ == "__main__":
    main()
