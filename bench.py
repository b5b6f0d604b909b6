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


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_coupled(mesh, params, state, mems, body_force=None):
    """The FP64 oracle of the coupled step (fluid + membranes + 4-point IBM)."""
    from lbfs_amd import abi
    from oracle import oracle as orc
    fl = orc.OracleFluid(mesh, orc.OR_D3Q19, visc=params.visc, dt=params.dt)
    fl.set_state(*state)
    mos = [orc.OracleMembrane(m.nodes, m.tri, m.params, nodes_stress_free=m.sf, nodes_curvature=m.cv) for m in mems]
    bf = tuple(params.body_force) if body_force is None else tuple(body_force)
    return fl, mos, orc.OracleCoupled(fl, mesh, mos, abi.IBM_COS4, body_force=bf)


def sample_errors(gpu, fl, mos):
    """Relative L2 of the GPU sample's velocity and node positions against
    the FP64 oracle's."""
    _, u, v, w = gpu["state"]
    _, uo, vo, wo = fl.state()
    out = {"velocity_rel_l2": rel_l2(np.concatenate([u, v, w]), np.concatenate([uo, vo, wo]))}
    if mos:
        out["nodes_rel_l2"] = rel_l2(gpu["nodes"], np.concatenate([m.X.reshape(-1, 3) for m in mos]))
    return out


def cpu_leg(mesh, params, state, mems, c2_sample, tgv, ens_sample):
    """The CPU baseline and the FP64 errors of the GPU samples.

    Times the FP64 oracle (one thread) on the headline workload's first
    c2_sample["steps"] coupled steps (the bounded sample) and on the whole
    Taylor-Green case (1000 steps, the CPU-runnable configuration), and
    compares each GPU sample taken on the same inputs with the oracle's result:
    the channel + cell, the Taylor-Green vortex and one instance of the
    ensemble run alone."""
    from oracle import oracle as orc
    steps = c2_sample["steps"]
    fl, mos, cp = oracle_coupled(mesh, params, state, mems)
    t0 = time.perf_counter()
    cp.step(steps)
    dt = time.perf_counter() - t0
    parity = {"channel_rbc": dict(sample_errors(c2_sample, fl, mos),
                                  sample="%d coupled steps from rest, FP32 GPU vs FP64 oracle" % steps)}
    c1 = None
    if tgv is not None:
        mt, pt, stt = tgv["case"]
        o = orc.OracleFluid(mt, orc.OR_D2Q9, visc=pt.visc, dt=pt.dt)
        o.set_state(*stt)
        t1 = time.perf_counter()
        o.step(tgv["steps"])
        dt1 = time.perf_counter() - t1
        c1 = {"workload": "64x64 D2Q9 Taylor-Green vortex, %d steps, FP64" % tgv["steps"],
              "seconds": dt1, "mlups": mt.n_cells * tgv["steps"] / dt1 / 1e6,
              "gpu_ms_total": tgv["gpu_ms"]}
        parity["taylor_green"] = dict(sample_errors({"state": tgv["state"]}, o, None),
                                      sample="%d steps, FP32 GPU vs FP64 oracle" % tgv["steps"])
    if ens_sample is not None:
        mi, bf, st_i, mm = ens_sample["instance"]
        efl, emos, ecp = oracle_coupled(mi, params, st_i, mm, body_force=bf)
        ecp.step(ens_sample["steps"])
        parity["ensemble_instance"] = dict(sample_errors(ens_sample, efl, emos), sample=ens_sample["label"])
    return {"value": mesh.n_cells * steps / dt / 1e6, "unit": "MLUPS", "cores": 1, "kind": "port",
            "sample": "%d coupled steps (fluid + 2562-node cell + 4-point IBM) of the 128x64x64 D3Q19 channel, "
                      "FP64 oracle, one thread, %.1f s" % (steps, dt),
            "host": host_cpu(), "c1_full": c1}, parity


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def profile_record(kernel, tile=None):
    """The committed rocprofv3 counter summary of `kernel` (a stage name) on
    this workload: the entry for the tile shape run ("fluid@16x8x4") where
    there is one, else the 8x8x8 entry, or None."""
    path = os.path.join(REPO, "profiles", "traffic_channel_128x64x64.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if tile is not None:
        rec = d.get("%s@%dx%dx%d" % (kernel, tile[0], tile[1], tile[2]))
        if rec is not None:
            return rec
    return d.get(kernel)


def measured_traffic(kernel, tile=None):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE per dispatch, each
    scaled by the same counter's reading on a copy of known size), or None."""
    rec = profile_record(kernel, tile)
    return None if rec is None else rec.get("bytes_per_launch")


# FP32 vector issue rate of the MI355X: 256 CUs x 4 SIMDs at 2.4 GHz, a
# wave64 VALU instruction occupying its 16-lane SIMD for 4 cycles (no packed
# FP32: the build turns it off)
VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4


def valu_roofline(kernel, launch_ms, tile=None):
    """The dominant kernel against its vector-ALU floor: wave-level VALU
    instructions per launch (SQ_INSTS_VALU from the committed counter summary)
    over the chip's issue rate, as a fraction of the measured launch time."""
    rec = profile_record(kernel, tile)
    insts = None if rec is None else rec.get("valu_insts_per_launch")
    if not insts:
        return None
    floor_ms = insts / VALU_WAVE_INSTS_PER_S * 1e3
    return {"insts_per_launch": insts, "peak_insts_per_s": VALU_WAVE_INSTS_PER_S, "floor_ms": floor_ms,
            "frac": floor_ms / launch_ms}


def timed(solver, steps, warmup):
    """Wall time per step of `steps` whole steps after `warmup`, synchronised
    on both sides (inputs resident in HBM)."""
    solver.step(warmup)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(steps)
    solver.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def new_solver(mesh, params, fused=None):
    """A solver with the given fluid pass (None: the library's default)."""
    from lbfs_amd import abi
    s = abi.Solver(mesh, params)
    if fused is not None:
        s.set_fused(fused)
    return s


def coupled_pass(args, mesh, params, state, mems, steps):
    """The fluid pass (fused, or the gradient and flux passes: the same bits)
    for a coupled step: --fused when given, else the faster of the two over
    `steps` whole coupled steps (the passes share the GPU with the coupling
    chain differently than in the plain step).  Returns (fused, {fused: ms})."""
    from lbfs_amd import abi
    if args.fused is not None:
        return bool(args.fused), {}
    times = {}
    for fz in (True, False):
        try:
            s = new_solver(mesh, params, fz)
        except abi.LbfsError:
            continue  # the tiles do not fit the fused pass
        try:
            s.set_state(*state)
            s.add_membranes(mems, ibm_kernel=abi.IBM_COS4)
            s.step(5)
            s.sync()
            times[fz] = s.time_steps(steps) / steps
        finally:
            s.close()
    return min(times, key=times.get), times


def pass_name(fz):
    return "library default" if fz is None else ("fused" if fz else "two-pass")


def run_suspension(args, params, log):
    """The 256-cell suspension (256x128x128, three resolutions, 20 %
    hematocrit): its step and its plain step on the same mesh, per kernel."""
    from lbfs_amd import abi, workloads
    log("building the 256-cell suspension")
    m3, p3, s3, mems, d0 = workloads.suspension()
    p3.tile[0], p3.tile[1], p3.tile[2] = params.tile[0], params.tile[1], params.tile[2]
    p3.block_threads = params.block_threads
    # the fluid pass (fused or two passes, the same bits) by the plain step of this mesh
    times = {}
    for fz in ((True, False) if args.fused is None else (bool(args.fused),)):
        try:
            b_plain = new_solver(m3, p3, fz)
        except abi.LbfsError:
            continue
        b_plain.set_state(*s3)
        times[fz] = timed(b_plain, max(args.steps // 4, 10), 5)
        b_plain.close()
    ms3_plain = min(times.values())
    fused, ctimes = coupled_pass(args, m3, p3, s3, mems, 10)
    t0 = time.perf_counter()
    b = new_solver(m3, p3, fused)
    t1 = time.perf_counter()
    b.set_state(*s3)
    b.add_membranes(mems, ibm_kernel=abi.IBM_COS4)
    setup = {"create_s": t1 - t0, "membranes_s": time.perf_counter() - t1}
    ms3 = timed(b, max(args.steps // 4, 10), 5)
    k3 = b.time_kernels(10)
    b.close()
    nodes = sum(len(mm.nodes) // 3 for mm in mems)
    log("suspension step %.4f ms (plain %.4f ms)" % (ms3, ms3_plain))
    return {"workload": "256x128x128 jittered hex, 256 Evans-Fung cells (d0 %.1f, %d nodes), 20%% hematocrit, "
                        "D3Q19" % (d0, nodes),
            "fluid_pass": pass_name(fused), "plain_fluid_pass": pass_name(min(times, key=times.get)),
            "ms_per_step": ms3, "plain_ms_per_step": ms3_plain, "ratio_vs_plain": ms3 / ms3_plain,
            "mlups": m3.n_cells / (ms3 * 1e-3) / 1e6, "kernels_ms": k3, "setup": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="coupled steps of the headline workload in the CPU baseline sample (about 1 s each)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", type=int, nargs=3, default=None,
                    help="fixed tile shape of the flux/update kernel (skips the autotuner)")
    ap.add_argument("--threads", type=int, default=0, help="workgroup size with --tile (256 or 512)")
    ap.add_argument("--fused", type=int, choices=(0, 1), default=None,
                    help="with --tile: 1 the fused fluid pass, 0 the gradient and flux passes (default: the library's)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device-to-device copies (the reference for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--long-steps", type=int, default=5000, help="steps of the long channel + RBC run (0: none)")
    ap.add_argument("--tgv-steps", type=int, default=1000, help="steps of the Taylor-Green check (0: none)")
    ap.add_argument("--only-rbc", action="store_true",
                    help="time the channel + RBC step alone (no plain step, suspension, ensemble or adjoint): "
                         "profiling runs")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-every", type=int, default=100)
    ap.add_argument("--only-batched", action="store_true",
                    help="time the 256-cell suspension alone (no channel, adjoint or CPU baseline): profiling runs")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-instance ensemble (edges from {32,48,64,96}, 1-8 cells of 642 nodes each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU design is documented, not built); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 1 << 30, "copies": 3}))
        return

    from lbfs_amd import abi, workloads
    log = lambda m: print("bench: " + m, file=sys.stderr, flush=True)  # noqa: E731

    if args.only_batched:
        p0 = abi.make_params()
        if args.tile:
            from lbfs_amd import autotune
            p0 = autotune.with_tile(p0, tuple(args.tile), args.threads or 256)
        print(json.dumps({"batched": run_suspension(args, p0, log)}))
        return

    # --- 128x64x64 channel with one 2562-node RBC (the headline step) ---
    log("building the 128x64x64 channel")
    mesh, params, state = workloads.channel()
    n = mesh.n_cells
    tuning = None
    fused = None  # the library's default fluid pass
    if args.tile:
        from lbfs_amd import autotune
        params = autotune.with_tile(params, tuple(args.tile), args.threads or 256)
        fused = None if args.fused is None else bool(args.fused)
    elif not args.no_autotune:
        # tile shape, workgroup size and fluid pass (results do not depend on
        # them); tuned on the coupled step this line reports (channel + RBC),
        # outside the timed runs, among the 512-thread workgroups: the
        # suspension and the adjoint reported beside it run on the same tiles,
        # and their tile spreading and face transpose are built for 8x8x8
        # tiles (on 256-cell tiles they run 20-25 % slower, and the adjoint's
        # backward some 5 % slower on 16x8x4, while the channel step differs
        # by under 1 %); a candidate displaces 8x8x8 (listed first) only when
        # over 1 % faster
        from lbfs_amd import autotune
        params, fused, table = autotune.tune(mesh, params, state=state, steps=50, warmup=10,
                                             candidates=[c for c in autotune.CANDIDATES if c[3] == 512],
                                             membranes=workloads.single_rbc(mesh), ibm_kernel=abi.IBM_COS4,
                                             margin=0.01)
        tuning = [{"tile": list(t), "threads": nt, "fused": f, "ms_per_step": ms} for t, nt, f, ms in table]
        log("autotuned tile %s, %d threads, %s fluid pass" % (list(params.tile), params.block_threads,
                                                              "fused" if fused else "two-pass"))
    if args.only_rbc:
        args.no_batched = args.no_adjoint = args.no_ensemble = True
        args.long_steps = args.tgv_steps = 0
        ms_plain = float("nan")
    else:
        plain = new_solver(mesh, params, fused)
        plain.set_state(*state)
        ms_plain = timed(plain, args.steps, args.warmup)
        plain.close()
        log("plain fluid step %.4f ms" % ms_plain)

    # the coupled step's own fluid pass (the plain step keeps the one tuned above)
    plain_fused = fused
    if not args.tile or args.fused is not None:
        fused, coupled_times = coupled_pass(args, mesh, params, state, workloads.single_rbc(mesh), 50)
        log("coupled step: %s fluid pass %s" % (pass_name(fused), {pass_name(k): round(v, 5)
                                                                   for k, v in coupled_times.items()}))
    t0 = time.perf_counter()
    solver = new_solver(mesh, params, fused)
    t1 = time.perf_counter()
    solver.set_state(*state)
    rbc = workloads.single_rbc(mesh)
    solver.add_membranes(rbc, ibm_kernel=abi.IBM_COS4)
    setup = {"create_s": t1 - t0, "membranes_s": time.perf_counter() - t1}
    # the sample the CPU baseline repeats in FP64: the first steps from rest
    c2_sample = None
    if not args.no_cpu_baseline:
        solver.step(args.cpu_steps)
        c2_sample = {"steps": args.cpu_steps, "state": solver.get_state(), "nodes": solver.membrane_nodes()[0]}
    ms_step = timed(solver, args.steps, args.warmup)
    mlups = n / (ms_step * 1e-3) / 1e6
    ev_ms = solver.time_steps(args.steps) / args.steps
    log("RBC step %.4f ms (events %.4f ms)" % (ms_step, ev_ms))
    kms = solver.time_kernels(min(args.steps, 50))
    names = list(kms.keys())
    mbytes = solver.model_bytes()
    step_bytes = float(np.sum(mbytes))
    dom = max(range(len(names)), key=lambda i: kms[names[i]])
    # the dominant kernel's duration: back-to-back launches between two HIP
    # events on the library stream (no per-launch event overhead)
    dom_ms = solver.time_stage(names[dom], 50) if names[dom] in ("gradients", "flux_update", "fluid") \
        else kms[names[dom]]
    achieved = mbytes[dom] / (dom_ms * 1e-3) / 1e9
    step_gbps = step_bytes / (ms_step * 1e-3) / 1e9
    diag = solver.diagnostics(ref_rho=1.0, ref_u=0.01)
    mglob = solver.membrane_globals()[0]
    solver.close()

    # --- the same channel + RBC over 5000 steps from rest: stability and
    # conservation of the long run (not timed against the others) ---
    long_run = None
    if args.long_steps > 0:
        log("long run: %d steps" % args.long_steps)
        lr = new_solver(mesh, params, fused)
        lr.set_state(*state)
        lr.add_membranes(workloads.single_rbc(mesh), ibm_kernel=abi.IBM_COS4)
        d0 = lr.diagnostics(ref_rho=1.0, ref_u=0.01)
        g0 = lr.membrane_globals()[0]
        t0 = time.perf_counter()
        lr.step(args.long_steps)
        lr.sync()
        t1 = time.perf_counter()
        d1 = lr.diagnostics(ref_rho=1.0, ref_u=0.01)
        g1 = lr.membrane_globals()[0]
        X, _, _ = lr.membrane_nodes()
        lr.close()
        long_run = {"steps": args.long_steps, "ms_per_step": (t1 - t0) * 1e3 / args.long_steps,
                    "mass_rel_change": d1["total_mass"] / d0["total_mass"] - 1.0, "max_speed": d1["max_speed"],
                    "membrane_volume_rel_change": g1[0] / g0[0] - 1.0, "membrane_area_rel_change": g1[1] / g0[1] - 1.0,
                    "membrane_centroid": [float(v) for v in X.mean(axis=0)], "finite": bool(np.isfinite(X).all())}
        log("long run: %.4f ms/step, volume change %.2e" % (long_run["ms_per_step"], long_run["membrane_volume_rel_change"]))

    # --- Taylor-Green vortex, 64x64 D2Q9 on a jittered periodic mesh (the
    # reference's 2D verification case): amplitude against exp(-2 nu k^2 t) ---
    taylor_green = None
    if args.tgv_steps > 0:
        mt, pt, stt = workloads.tgv_2d()
        tg = abi.Solver(mt, pt)
        tg.set_state(*stt)
        tg.sync()
        t0 = time.perf_counter()
        tg.step(args.tgv_steps)
        tg.sync()
        tgv_ms = (time.perf_counter() - t0) * 1e3
        tgv_state = tg.get_state()
        _, u1, v1, _ = tgv_state
        tg.close()
        k = 2 * np.pi / 64
        tt = args.tgv_steps * pt.dt
        decay = float(np.exp(-2 * pt.visc * k * k * tt))
        amp = float(np.dot(np.concatenate([u1, v1]), np.concatenate([stt[1], stt[2]])) /
                    np.dot(np.concatenate([stt[1], stt[2]]), np.concatenate([stt[1], stt[2]])))
        taylor_green = {"workload": "64x64 D2Q9, jitter 0.05, periodic, Re 100, U0 0.01, dt %g, %d steps"
                                    % (pt.dt, args.tgv_steps),
                        "precision": "FP32 on the GPU (the configuration names FP64: its FP64 error is in "
                                     "parity_fp64.taylor_green)",
                        "amplitude": amp, "analytic": decay, "rel_error": amp / decay - 1.0, "gpu_ms_total": tgv_ms}
        log("Taylor-Green: amplitude %.6f, analytic %.6f" % (amp, decay))

    # --- 256x128x128 suspension of 256 RBCs of mixed resolution ---
    batched = None
    if not args.no_batched:
        batched = run_suspension(args, params, log)

    # --- adjoint of the channel + RBC step, node loss against a run at
    # moduli x 1.1: every step recorded in HBM (no replay), and checkpoints
    # every K steps with segment replays (little memory) ---
    adjoint = adjoint_ckpt = None
    if not args.no_adjoint:
        def run_adjoint(every):
            log("adjoint: %d steps, checkpoints every %d" % (args.adjoint_steps, every))
            mems = workloads.single_rbc(mesh)
            mp = mems[0].params
            mu, kl, kg = mp["shear_modulus"], mp["local_bending_modulus"], mp["global_bending_modulus"]
            a = new_solver(mesh, params, fused)
            try:
                a.set_state(*state)
                a.add_membranes(mems, ibm_kernel=abi.IBM_COS4)
                a.set_membrane_moduli(0, 1.1 * mu, 1.1 * kl, 1.1 * kg)
                a.adjoint_forward(args.adjoint_steps, every)
                target, _, _ = a.membrane_nodes()
                a.adjoint_rewind()
                a.set_membrane_moduli(0, mu, kl, kg)
                a.sync()
                t0 = time.perf_counter()
                a.adjoint_forward(args.adjoint_steps, every)
                t1 = time.perf_counter()
                X, _, _ = a.membrane_nodes()
                g = a.adjoint_backward(lam_nodes=2.0 * (X - target))
                t2 = time.perf_counter()
            finally:
                a.close()
            fwd = (t1 - t0) * 1e3 / args.adjoint_steps
            bwd = (t2 - t1) * 1e3 / args.adjoint_steps
            log("adjoint: forward %.4f ms/step, backward %.4f ms/step" % (fwd, bwd))
            return {"workload": "128x64x64 channel + RBC, %d steps, checkpoints every %d, "
                                "L = sum |X - X(moduli x 1.1)|^2" % (args.adjoint_steps, every),
                    "loss": float(np.sum((X - target) ** 2)),
                    "grad": {"shear_modulus": float(g["moduli"][0, 0]),
                             "local_bending_modulus": float(g["moduli"][0, 1]),
                             "global_bending_modulus": float(g["moduli"][0, 2]),
                             "body_force_x": float(g["body_force"][0])},
                    "forward_ms_per_step": fwd, "backward_ms_per_step": bwd,
                    "backward_vs_forward": bwd / fwd, "backward_vs_graph_step": bwd / ms_step,
                    "note": ("forward records every step; backward = reverse kernels"
                             if every >= args.adjoint_steps else
                             "forward records the last segment; backward = segment replays + reverse kernels")}
        # every step recorded (about 77 MB per step on this mesh): if the device
        # cannot hold the whole run, record the last quarter or less instead
        adjoint = None
        for every in (args.adjoint_steps, args.adjoint_steps // 4, args.adjoint_every):
            try:
                adjoint = run_adjoint(every)
                break
            except abi.LbfsError as e:
                log("adjoint with %d recorded steps failed (%s); fewer records" % (every, e))
        adjoint_ckpt = run_adjoint(args.adjoint_every)

    ensemble = ens_sample = None
    if not args.no_ensemble:
        log("building the 64-instance ensemble")
        m5, p5, s5, mems5 = workloads.ensemble()
        p5.tile[0], p5.tile[1], p5.tile[2] = params.tile[0], params.tile[1], params.tile[2]
        p5.block_threads = params.block_threads
        e_plain = new_solver(m5, p5, plain_fused)
        e_plain.set_state(*s5)
        ms5_plain = timed(e_plain, max(args.steps // 4, 10), 5)
        e_plain.close()
        fused5, _ = coupled_pass(args, m5, p5, s5, mems5, 10)
        t0 = time.perf_counter()
        e = new_solver(m5, p5, fused5)
        t1 = time.perf_counter()
        e.set_state(*s5)
        e.add_membranes(mems5, ibm_kernel=abi.IBM_COS4)
        setup5 = {"create_s": t1 - t0, "membranes_s": time.perf_counter() - t1}
        if not args.no_cpu_baseline:
            # the instance with the most cells, its first 10 steps from rest
            counts = np.bincount([mm.block for mm in mems5], minlength=len(m5.members))
            pick = int(np.argmax(counts))
            e.step(10)
            X5 = e.membrane_nodes()[0]
            starts = np.cumsum([0] + [len(mm.nodes) // 3 for mm in mems5])
            mine = [k for k, mm in enumerate(mems5) if mm.block == pick]
            mi = m5.members[pick]
            ens_sample = {"steps": 10, "state": [m5.split(x)[pick] for x in e.get_state()],
                          "nodes": np.concatenate([X5[starts[k]:starts[k + 1]] for k in mine]),
                          "instance": (mi, m5.body_forces[3 * pick: 3 * pick + 3],
                                       [m5.split(x)[pick] for x in s5],
                                       [mems5[k] for k in mine]),
                          "label": "instance %d (%dx%dx%d, %d cells), 10 steps from rest inside the ensemble, "
                                   "FP32 GPU vs the FP64 oracle on the instance alone"
                                   % (pick, mi.nx, mi.ny, mi.nz, len(mine))}
        ms5 = timed(e, max(args.steps // 4, 10), 5)
        k5 = e.time_kernels(10)
        e.close()
        ensemble = {"workload": "64 independent jittered channel boxes (edges from {32,48,64,96}, seeds 5000+i, "
                                "each with its own duct body force), %d cells of 642 nodes, D3Q19, one mesh"
                                % len(mems5),
                    "cells": m5.n_cells, "fluid_pass": pass_name(fused5), "plain_fluid_pass": pass_name(plain_fused),
                    "ms_per_step": ms5, "plain_ms_per_step": ms5_plain,
                    "ratio_vs_plain": ms5 / ms5_plain, "mlups": m5.n_cells / (ms5 * 1e-3) / 1e6, "kernels_ms": k5,
                    "setup": setup5}
        log("ensemble step %.4f ms (plain %.4f ms)" % (ms5, ms5_plain))

    cpu = parity = None
    if not args.no_cpu_baseline:
        log("CPU baseline (FP64 oracle, %d coupled steps) and FP64 errors of the GPU samples" % args.cpu_steps)
        tgv = None if taylor_green is None else {"case": (mt, pt, stt), "steps": args.tgv_steps,
                                                 "state": tgv_state, "gpu_ms": tgv_ms}
        cpu, parity = cpu_leg(mesh, params, state, rbc, c2_sample, tgv, ens_sample)

    out = {
        "metric": METRIC,
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered hex mesh, rest state + body force, seeded RBC shapes)",
        "config": {"workload": "channel 128x64x64 jittered hex (+-10%, seed 2), x-periodic, walls y/z, D3Q19, "
                               "body-force Poiseuille, one Evans-Fung RBC 2562 nodes / 5120 triangles, 4-point IBM",
                   "cells": n, "faces": mesh.n_faces, "lattice": "D3Q19",
                   "tile": [params.tile[0], params.tile[1], params.tile[2]],
                   "block_threads": params.block_threads or 256,
                   "fluid_pass": pass_name(fused), "plain_fluid_pass": pass_name(plain_fused)},
        "autotune": tuning,
        "achieved_hbm_gbps_model": step_gbps,
        "plain_ms_per_step": None if args.only_rbc else ms_plain,
        "plain_mlups": None if args.only_rbc else n / (ms_plain * 1e-3) / 1e6,
        "rbc_step_vs_plain": None if args.only_rbc else ms_step / ms_plain,
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(names[dom], tuple(params.tile)),
                     "model_bytes_per_launch": mbytes[dom], "avg_launch_ms": dom_ms,
                     "valu": valu_roofline(names[dom], dom_ms, tuple(params.tile))},
        "kernels_ms": kms,
        "setup": setup,
        "event_ms_per_step": ev_ms,
        "max_speed": diag["max_speed"],
        "membrane_volume_rel_change": mglob[0] / mglob[3] - 1.0,
        "long_run": long_run,
        "taylor_green": taylor_green,
        "batched": batched,
        "ensemble": ensemble,
        "adjoint": adjoint,
        "adjoint_checkpointed": adjoint_ckpt,
        "cpu_baseline": cpu,
        "parity_fp64": parity,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
