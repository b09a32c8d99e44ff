#!/usr/bin/env python3
"""Benchmark: RK4 DOF-updates/s of the DG acoustic solver core on one MI355X.

Workload (the 3D APE case the headline metric is quoted on): cube [-1,1]^3, 32^3 hexes x 6
Kuhn tetrahedra (196,608 elements), interior vertices jittered by U(+-0.15h) (seed 2), p=3
(Np=20), 4 fields (p, vx, vy, vz), mean flow (30,0,0), rho0=1.225, c0=340, monopole at the
origin (size 0.1, amplitude 0.1, 200 Hz), all boundaries absorbing, dt = r_min /
((c0+|v0|)(2p+1)) (about half the RK4 stability limit), two observers.

Each precision first runs a fixed accuracy leg of 2000 RK4 steps from a zero state (the
metric's condition: the pressure field and the observer pressure series of the mixed path
within 1e-6 relative L2 of FP64 after 2000 steps, whatever --steps is), then --warmup and the
--steps timed steps continue from the state that leg ends in (a filled field, not zeros). The
mixed-precision path (FP32 state and fluxes, compensated RK4 accumulation) is `value`. A
DOF-update is one DOF advanced one full RK4 step; each step includes the reference's per-step
work (source overwrite, residuals, observers). The stage kernels are timed afterwards with HIP
events (dg_profile_steps, which advances the state further; nothing reads it after that).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("dgfem-acoustic_amd",):
    sys.path.insert(0, os.path.join(ROOT, sub))

METRIC = "RK4 DOF-updates/s, mixed-precision vs FP64 on 1 GPU at ≤1e-6 rel L2 pressure"
HBM_PEAK_GBPS = 8000.0
# Algorithmic HBM bytes per DOF-update (classic 3-register RK4: 17 FP64 / 19 FP32 state
# passes, neighbour face traces once per stage, per-element geometry once per stage).
BYTES_PER_DOF = {"fp64": 212.0, "mixed": 115.0}


def copy_bandwidth(device=0, nbytes=2 << 30, reps=10):
    """On-box HBM reference: GB/s (read + write) of a streaming 16-byte copy kernel
    (dgfem-acoustic_amd/libdgfem_probe.so), or None if it is not built."""
    import ctypes as C
    path = os.path.join(ROOT, "dgfem-acoustic_amd", "libdgfem_probe.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.dgprobe_copy_gbps.restype = C.c_double
    lib.dgprobe_copy_gbps.argtypes = [C.c_int32, C.c_int64, C.c_int32]
    v = lib.dgprobe_copy_gbps(device, nbytes, reps)
    return v if v > 0 else None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH/WRITE summary."""
    path = os.path.join(ROOT, "profiles", "stage_kernel_c2_traffic.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch " + kernel + ":"):
                return float(line.split(":")[-1])
    except OSError:
        pass
    return None


def build_case(n, seed=2):
    import dgfem_amd as dg

    m = dg.box_mesh(3, n, 3, jitter=0.15, seed=seed)
    bc = dg.resolve_bc_sides(m, [(0x3F, dg.DG_BC_ABSORBING)])
    return m, bc


def make_solver(m, bc, precision, steps_per_graph):
    import dgfem_amd as dg

    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=1.225, c0=340.0,
                  precision=precision, use_graph=True, steps_per_graph=steps_per_graph)
    s.add_monopole(0.0, 0.0, 0.0, 0.1, 0.1, 200.0, 0.0, 1e9)
    s.add_observer(0.5, 0.0, 0.0, 0.1)
    s.add_observer(0.0, -0.5, 0.25, 0.1)
    return s


def accuracy_then_timed_run(s, acc_steps, steps, warmup, dt):
    """acc_steps steps from a zero state (state and observer series kept), then warmup and the
    timed steps from where they end (the source clock continues)."""
    s.set_state(np.zeros(s.state_shape))
    r_acc = s.run_steps(acc_steps, dt)
    u_acc = s.get_state()
    t = acc_steps * dt
    if warmup > 0:
        s.run_steps(warmup, dt, t_start=t)
        t += warmup * dt
    t0 = time.perf_counter()
    r = s.run_steps(steps, dt, t_start=t)
    wall = time.perf_counter() - t0
    return (u_acc, r_acc.observers), r, wall


def cpu_baseline(m, bc, dt, budget_s):
    """Oracle (the reference algorithm restated, OpenMP) on the same mesh: one RHS evaluation
    (one RK4 stage) per sample, repeated while within the time budget; a step is 4 stages."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc

    groups = [(0x3F, 0)]
    gn = orc.group_node_sets(m.element_vertices, m.face_side, groups)
    o = orc.Oracle(m.node_coords, m.element_vertices, 3, gn, [0], v0=(30.0, 0.0, 0.0), rho0=1.225, c0=340.0)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    rng = np.random.default_rng(0)
    u = rng.normal(size=(4, m.K * m.Np)) * 1e-3
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        o.rhs(u, dt, threads)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start + times[-1] > budget_s:
            break
    per_stage = float(np.median(times))
    # one single-threaded evaluation beside it
    t0 = time.perf_counter()
    o.rhs(u, dt, 1)
    one = time.perf_counter() - t0
    dof = m.K * m.Np * 4
    return {"value": dof / (4.0 * per_stage), "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "value_1_core": dof / (4.0 * one), "cpu_model": cpu_model(),
            "sample": f"{len(times)} RHS evaluation(s) (one RK4 stage each) of the oracle on the full mesh "
                      f"({m.K} tets, p=3), median {per_stage:.2f} s per stage, x4 per step; oracle quadrature "
                      f"{o.num_volume_points} points per element; one more evaluation on 1 core: {one:.2f} s"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=32, help="hexes per side")
    ap.add_argument("--steps-per-graph", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=25.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--mixed-kernel", default=None,
                    help="stage kernel of the mixed path (DGFEM_KERNEL value; default: the library's choice)")
    ap.add_argument("--fp64-kernel", default=None,
                    help="stage kernel of the FP64 path (DGFEM_KERNEL value; default: the library's choice)")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="steps of the mixed-vs-FP64 accuracy leg (the metric's condition is at 2000)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import dgfem_amd as dg

    m, bc = build_case(args.n)
    probe = make_solver(m, bc, dg.DG_FP64, args.steps_per_graph)
    info = probe.info()
    c0, vmag, p = 340.0, 30.0, 3
    dt = info.min_inradius / ((c0 + vmag) * (2 * p + 1))
    probe.close()
    dof = m.K * m.Np * 4
    print(f"# workload: K={m.K} Np={m.Np} DOF={dof} dt={dt:.4e} accuracy steps={args.accuracy_steps} "
          f"timed steps={args.steps}", file=sys.stderr, flush=True)

    out = {}
    states = {}
    kernels = {}
    for name, prec in (("fp64", dg.DG_FP64), ("mixed", dg.DG_MIXED)):
        want = args.mixed_kernel if name == "mixed" else args.fp64_kernel
        if want:
            os.environ["DGFEM_KERNEL"] = want
        s = make_solver(m, bc, prec, args.steps_per_graph)
        if want:
            os.environ.pop("DGFEM_KERNEL", None)
        kernels[name] = dg.KERNEL_NAMES[s.info().stage_kernel]
        states[name], r, wall = accuracy_then_timed_run(s, args.accuracy_steps, args.steps, args.warmup, dt)
        ms = s.profile_steps(dt, args.profile_steps)
        stage_ms = ms[0]
        out[name] = {"wall_s": wall, "device_ms": r.device_ms, "ms_per_step": 1e3 * wall / args.steps,
                     "value": dof * args.steps / wall, "stage_ms": stage_ms, "source_ms": ms[1], "post_ms": ms[2]}
        print(f"# {name}: {out[name]['value']:.4e} DOF-updates/s, {out[name]['ms_per_step']:.4f} ms/step, "
              f"stage kernels {['%.4f' % x for x in stage_ms]} ms, post {ms[2]:.4f} ms", file=sys.stderr, flush=True)
        s.close()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    (u64, o64), (umx, omx) = states["fp64"], states["mixed"]
    accuracy = {"steps": args.accuracy_steps, "from": "zero state, monopole at the origin",
                "rel_l2_pressure": rel(umx[0], u64[0]), "rel_l2_observer_p": rel(omx[..., 0], o64[..., 0]),
                "rel_l2_velocity": [rel(umx[f], u64[f]) for f in (1, 2, 3)],
                "bar": 1e-6}
    accuracy["meets_bar"] = bool(accuracy["rel_l2_pressure"] <= 1e-6 and accuracy["rel_l2_observer_p"] <= 1e-6)

    mixed = out["mixed"]
    kernel_name = {"valu": "stageKernel<3,3,4,float,true>",
                   "vm": "stageKernel<3,3,4,float,true,VM>",
                   "wave": "stageKernelWave<3,3,float,true>",
                   "mm": "stageKernelMM<3,3,float,true>",
                   "bf16x6": "stageKernelBf16<3,3,4,3>",
                   "bf16x3": "stageKernelBf16<3,3,4,2>"}[kernels["mixed"]]
    stage_sum_ms = sum(mixed["stage_ms"])
    achieved = BYTES_PER_DOF["mixed"] * dof / (stage_sum_ms * 1e-3) / 1e9
    achieved_wall = BYTES_PER_DOF["mixed"] * dof / (mixed["ms_per_step"] * 1e-3) / 1e9
    stage_sum64 = sum(out["fp64"]["stage_ms"])
    achieved64 = BYTES_PER_DOF["fp64"] * dof / (stage_sum64 * 1e-3) / 1e9
    achieved64_wall = BYTES_PER_DOF["fp64"] * dof / (out["fp64"]["ms_per_step"] * 1e-3) / 1e9

    copy_gbps = copy_bandwidth()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(m, bc, dt, args.cpu_budget)

    line = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3D APE, [-1,1]^3 32^3x6 jittered Kuhn tets (K=196608), p=3, 4 fields, "
                               "v0=(30,0,0), rho0=1.225, c0=340, monopole 200 Hz at origin, absorbing, "
                               "CFL-limited dt",
                   "elements": m.K, "order": 3, "dofs": dof, "dt": dt, "seed": 2,
                   "precision": "mixed (FP32 state/fluxes, compensated RK4 accumulation)"},
        "fp64": {"value": out["fp64"]["value"], "ms_per_step": out["fp64"]["ms_per_step"],
                 "stage_ms": out["fp64"]["stage_ms"], "kernel": kernels["fp64"],
                 "roofline_frac": achieved64 / HBM_PEAK_GBPS,
                 "roofline_frac_wall": achieved64_wall / HBM_PEAK_GBPS},
        "speedup_vs_fp64": mixed["value"] / out["fp64"]["value"],
        "rel_l2_pressure_vs_fp64": accuracy["rel_l2_pressure"],
        "rel_l2_observer_p_vs_fp64": accuracy["rel_l2_observer_p"],
        "accuracy": accuracy,
        "timed_from": f"the state after the {args.accuracy_steps}-step accuracy leg and {args.warmup} warm-up steps",
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "achieved_wall": achieved_wall, "frac_wall": achieved_wall / HBM_PEAK_GBPS,
                     "copy_gbps": copy_gbps,
                     "frac_of_copy": achieved / copy_gbps if copy_gbps else None,
                     "traffic": measured_traffic(kernel_name) if args.n == 32 else None,
                     "kernel": kernel_name,
                     "stage_ms": mixed["stage_ms"],
                     "bytes_per_dof_update": BYTES_PER_DOF["mixed"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
