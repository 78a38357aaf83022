"""Benchmark of the implicit PIC step on one MI355X.

Metric: particle-steps/s including every deposit (one particle advanced one full
implicit step: moments, field solve, push, re-sort and post-push density), plus the
GMRES time per Maxwell solve. Workload: the one-rank production box, 128^3 cells,
64 particles per cell per species (electron-proton, m_p/m_e = 100, kT/mc^2 = 0.01,
omega_ce/omega_pe = 0.5, dt = 0.1/omega_pe), synthetic seeded initial state generated
on the GPU.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 128] [--ppc 64]
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=64)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--stop-rule", default="reference", choices=["reference", "true_residual"])
    ap.add_argument("--ortho", default="mgs", choices=["mgs", "cgs2"])
    ap.add_argument("--solver", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU baseline sample")
    return ap.parse_args()


def make_state_on_device(w, torch, device):
    """Seeded synthetic state on the GPU: uniform positions in each cell, Maxwell-Juettner
    momenta (tabulated inverse CDF, as picpp_amd.workloads does on the host)."""
    from picpp_amd import workloads as wl

    gen = torch.Generator(device=device)
    gen.manual_seed(w.seed)
    n = w.n
    ncell = n[0] * n[1] * n[2]
    xs, ps, counts = [], [], []
    for s in w.species:
        cnt = ncell * s.ppc
        cid = torch.arange(ncell, device=device, dtype=torch.int64).repeat_interleave(s.ppc)
        ci = torch.stack([cid // (n[1] * n[2]), (cid // n[2]) % n[1], cid % n[2]]).to(torch.float64)
        del cid
        lo = torch.tensor([v * w.dx for v in n], device=device, dtype=torch.float64)[:, None]
        x = lo + (ci + torch.rand((3, cnt), generator=gen, device=device, dtype=torch.float64)) * w.dx
        del ci
        ut, cdf = wl._mj_table(s.temperature)
        ut = torch.tensor(ut, device=device)
        cdf = torch.tensor(cdf, device=device)
        r = torch.rand(cnt, generator=gen, device=device, dtype=torch.float64)
        k = torch.searchsorted(cdf, r).clamp(1, len(cdf) - 1)
        t = ((r - cdf[k - 1]) / (cdf[k] - cdf[k - 1]).clamp_min(1e-300)).clamp(0, 1)
        mag = ut[k - 1] + t * (ut[k] - ut[k - 1])
        del r, k, t
        cost = 2 * torch.rand(cnt, generator=gen, device=device, dtype=torch.float64) - 1
        phi = 2 * math.pi * torch.rand(cnt, generator=gen, device=device, dtype=torch.float64)
        sint = torch.sqrt((1 - cost * cost).clamp_min(0))
        p = torch.stack([mag * sint * torch.cos(phi), mag * sint * torch.sin(phi), mag * cost]) * s.mass
        del cost, phi, sint, mag
        xs.append(x)
        ps.append(p)
        counts.append(cnt)
    x = torch.cat(xs, 1).contiguous()
    p = torch.cat(ps, 1).contiguous()
    return x, p, counts


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
    return {"value": npart * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle FP64 step, 32x32x8 cells x 40 ppc (%d particles), %d steps in %.1f s" % (npart, steps, el)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    from picpp_amd import workloads as wl
    from picpp_amd.simulation import Simulation

    dev = torch.device("cuda:0")
    n = (args.cells,) * 3
    w = wl.config2(n=n, ppc=args.ppc, steps=args.steps)
    stream = torch.cuda.current_stream(dev)
    sim = Simulation(w, precision=args.precision, stop_rule=args.stop_rule, ortho=args.ortho,
                     capacity=w.particles_total, stream=stream.cuda_stream, solver=args.solver)
    # fields: E = 0, B = B0 x (uniform); ghosted component planes
    N = w.N
    E = np.zeros((N[0] + 1, N[1] + 1, N[2] + 1, 3))
    B = np.zeros(N + (3,))
    B[...] = np.array(w.b0)
    st = wl.InitialState(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros((0, 3)), E, B)
    sim.load_state(st)
    x, p, counts = make_state_on_device(w, torch, dev)
    torch.cuda.synchronize()
    cnt = (ctypes.c_longlong * len(counts))(*counts)
    _D = ctypes.POINTER(ctypes.c_double)
    rc = sim.lib.pic_upload_particles_device(sim.h, cnt, ctypes.cast(x.data_ptr(), _D), ctypes.cast(p.data_ptr(), _D),
                                             None)
    sim._chk(rc, "pic_upload_particles_device")
    del x, p
    torch.cuda.empty_cache()
    npart = sim.num_particles()
    print("bench: %s, %d particles, precision %s" % (w.name, npart, args.precision), file=sys.stderr, flush=True)
    sim.initialize()
    for i in range(args.warmup):
        sim.step()
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    sim._chk(sim.lib.pic_profile_enable(sim.h, 1), "pic_profile_enable")
    solves = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        solves.append(sim.step())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    for i, s in enumerate(solves):
        print("step %d: gmres %d iterations, %.3f ms" % (i, s["iterations"], 1e3 * s["seconds"]), file=sys.stderr)
    # per-kernel statistics over the timed region (HIP events on the bench stream)
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
    dom = max(kernels, key=lambda q: kernels[q]["ms_total"])
    kd = kernels[dom]
    avg_ms = kd["ms_total"] / kd["launches"]
    per_launch = kd["bytes_total"] / kd["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    for nm, kv in sorted(kernels.items(), key=lambda q: -q[1]["ms_total"]):
        bw = kv["bytes_total"] / (kv["ms_total"] * 1e-3) / 1e9 if kv["ms_total"] > 0 else 0
        print("kernel %-18s %8.3f ms/step  %5d launches  %7.1f GB/s algorithmic" %
              (nm, kv["ms_total"] / args.steps, kv["launches"], bw), file=sys.stderr)
    gm_ms = 1e3 * float(np.mean([s["seconds"] for s in solves]))
    gm_it = float(np.mean([s["iterations"] for s in solves]))
    energy = sim.updateEnergy()
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
                               "%s GMRES, %s stop rule, %s orthogonalisation" % (
                                   n[0], n[1], n[2], args.ppc, npart, args.precision, args.solver, args.stop_rule,
                                   args.ortho),
                   "cells": list(n), "particles": npart, "precision": args.precision, "solver": args.solver},
        "gmres": {"ms_per_solve": gm_ms, "iterations_per_solve": gm_it, "stop_rule": args.stop_rule,
                  "solver": args.solver, "learned_preconditioner": False,
                  "true_residual": float(np.mean([s["true_residual"] for s in solves]))},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": per_launch},
        "kernels_ms_per_step": {nm: kv["ms_total"] / args.steps for nm, kv in kernels.items()},
        "energy": {"particles": energy[0], "E": energy[1], "B": energy[2], "total": energy[3]},
    }
    sim.close()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
