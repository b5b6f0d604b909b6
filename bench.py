#!/usr/bin/env python3
"""Benchmark of the per-timestep hot path on one MI355X.

Workload (one "step" = one full time step of the hot path over the mesh):
128x64x64 jittered hex channel (+-10 %, seed 2) as unstructured face/cell
connectivity, x-periodic, no-slip walls in y and z, D3Q19 FP32, body-force
Poiseuille flow.  Prints one JSON line with MLUPS (value), the modelled
HBM bandwidth of the step, the roofline of the dominant kernel (modelled bytes
per launch / average launch time from HIP events on the library's stream) and
a CPU baseline (the FP64 oracle on a bounded sample of the same workload).
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "lbfs-ibm-ade-sp_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "MLUPS + achieved HBM GB/s, 1 GPU; batched-RBC and adjoint step vs plain"
HBM_PEAK_GBPS = 8000.0


def cpu_baseline(mesh, params, steps):
    """FP64 oracle (single thread) on `steps` steps of the same mesh."""
    from oracle import oracle as orc
    o = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=params.visc, dt=params.dt)
    n = mesh.n_cells
    o.set_state(np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    force = np.tile(np.array([params.body_force[0], params.body_force[1], params.body_force[2]]), n)
    t0 = time.perf_counter()
    o.step(steps, force=force)
    dt = time.perf_counter() - t0
    return n * steps / dt / 1e6, dt


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU design is documented, not built); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from lbfs_amd import abi, workloads
    print("bench: building workload", file=sys.stderr, flush=True)
    mesh, params, state = workloads.channel()
    solver = abi.Solver(mesh, params)
    solver.set_state(*state)
    n = mesh.n_cells
    print("bench: warmup %d steps" % args.warmup, file=sys.stderr, flush=True)
    solver.step(args.warmup)
    solver.sync()

    # timed region: K whole steps, inputs resident in HBM, synchronised on both sides
    solver.sync()
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.sync()
    wall = time.perf_counter() - t0
    ev_ms = solver.time_steps(args.steps)
    ms_per_step = wall * 1e3 / args.steps
    mlups = n * args.steps / wall / 1e6
    print("bench: %d steps in %.3f s (events %.3f ms/step)" % (args.steps, wall, ev_ms / args.steps),
          file=sys.stderr, flush=True)

    # dominant kernel: modelled bytes per launch / average launch time
    kms = solver.time_kernels(min(args.steps, 50))
    names = list(kms.keys())
    mbytes = solver.model_bytes()
    step_bytes = float(np.sum(mbytes))
    dom = max(range(len(names)), key=lambda i: kms[names[i]])
    achieved = mbytes[dom] / (kms[names[dom]] * 1e-3) / 1e9
    step_gbps = step_bytes / (ms_per_step * 1e-3) / 1e9
    diag = solver.diagnostics(ref_rho=1.0, ref_u=0.01)
    solver.close()

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (FP64 oracle, %d steps)" % args.cpu_steps, file=sys.stderr, flush=True)
        cpu_mlups, cpu_s = cpu_baseline(mesh, params, args.cpu_steps)
        cpu = {"value": cpu_mlups, "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": "%d fluid steps of the same 128x64x64 D3Q19 mesh, FP64, single thread, %.1f s, host %s"
                         % (args.cpu_steps, cpu_s, cpu_model())}

    out = {
        "metric": METRIC,
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered hex mesh, rest state + body force)",
        "config": {"workload": "channel 128x64x64 jittered hex (+-10%, seed 2), x-periodic, walls y/z, "
                               "D3Q19, body-force Poiseuille, fluid step",
                   "cells": n, "faces": mesh.n_faces, "lattice": "D3Q19"},
        "achieved_hbm_gbps_model": step_gbps,
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "model_bytes_per_launch": mbytes[dom], "avg_launch_ms": kms[names[dom]]},
        "kernels_ms": kms,
        "event_ms_per_step": ev_ms / args.steps,
        "max_speed": diag["max_speed"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
