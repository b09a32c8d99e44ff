"""Throughput and accuracy of the GPU path on the named synthetic workloads (FP64 and mixed
precision from the same initial state; mixed-precision error reported against FP64).

  dense3d: 3D LEE, [-1,1]^3 as 44^3 hexes x 6 Kuhn tets (511,104 tets), interior vertices
           jittered by 0.15h, p=5 (56 nodes), 5 fields (p, v, rho), v0=(30,0,0), rho0=1.225,
           c0=340; sum of 8 seeded Gaussian pulses (centres U[-0.6,0.6]^3, widths U[0.05,0.15]
           used as exp(-r^2 / width^2), amplitudes N(0,1)); reflecting at z=-1 (group 2),
           absorbing elsewhere (group 1); CFL-limited dt.
  lee2d:   the 2D LEE case on [-1,1]^2 (96x96x2 triangles, p=3, v0=(-30,30), c0=100,
           reflecting, Gaussian pulse exp(-r^2/0.1)), CFL-limited dt.
  closure3d: throughput of the learned absorbing closure at the size of the 3D closure
           configuration's truncated domain: [-1,1]^3 as 64^3 x 6 jittered Kuhn tets (1,572,864
           tets, 49,152 boundary faces), p=3, APE, mixed precision, v0=(30,0,0), a monopole and
           8 observers; every boundary face absorbing (the reference's characteristic BC) or
           learned: the zero network (the exact characteristic condition) and a seeded network
           with a nonzero output layer (timing only, not trained).
  closure2d: the same at the 2D configuration's truncated size: [-1,1]^2 as 100^2 x 2 jittered
           triangles (20,000 triangles, 400 boundary faces), p=4.
  box3d:   [-1,1]^3 as n^3 x 6 jittered Kuhn tets at a given order (--n, --order), 4 fields,
           absorbing, a Gaussian pulse: stage-kernel comparisons across orders.

    python dgfem-acoustic_amd/workloads.py --workload dense3d --steps 200
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

import dgfem_amd as dg

RHO0 = 1.225


def dense3d(n=44, seed=3):
    m = dg.box_mesh(3, n, 5, jitter=0.15, seed=seed)
    # group 1 (absorbing): all sides but z-; group 2 (reflecting): z-  (ascending tag order)
    bc = dg.resolve_bc_sides(m, [(0x2F, dg.DG_BC_ABSORBING), (0x10, dg.DG_BC_REFLECTING)])
    rng = np.random.default_rng(seed)
    X = m.node_coords.reshape(-1, 3)
    p = np.zeros(len(X))
    for _ in range(8):
        c = rng.uniform(-0.6, 0.6, 3)
        w = rng.uniform(0.05, 0.15)
        a = rng.normal(0.0, 1.0)
        p += a * np.exp(-((X - c) ** 2).sum(axis=1) / (w * w))
    c0 = 340.0
    u = np.zeros((5, len(X)))
    u[0] = p
    u[4] = p / (c0 * c0)
    return dict(mesh=m, bc=bc, order=5, v0=(30.0, 0.0, 0.0), c0=c0, fields=dg.DG_FIELDS_LEE5, u0=u,
                observers=[(0.5, 0.0, 0.0, 0.05), (0.0, -0.5, 0.25, 0.05), (-0.3, 0.3, -0.9, 0.05)],
                describe=f"3D LEE, {n}^3x6 jittered Kuhn tets (K={m.K}), p=5, 5 fields, 8 Gaussian pulses, "
                         "reflecting z=-1 + absorbing, CFL-limited dt")


def lee2d(n=96, seed=1):
    m = dg.box_mesh(2, n, 3, jitter=0.2, seed=seed)
    bc = dg.resolve_bc_sides(m, [(0xF, dg.DG_BC_REFLECTING)])
    X = m.node_coords.reshape(-1, 3)
    u = np.zeros((4, len(X)))
    u[0] = np.exp(-(X ** 2).sum(axis=1) / 0.1)
    return dict(mesh=m, bc=bc, order=3, v0=(-30.0, 30.0, 0.0), c0=100.0, fields=dg.DG_FIELDS_REF4, u0=u,
                describe=f"2D LEE, {n}^2x2 jittered triangles (K={m.K}), p=3, reflecting, Gaussian pulse, "
                         "CFL-limited dt")


def box3d(n=32, order=3, seed=2):
    """[-1,1]^3 as n^3 x 6 jittered Kuhn tets at the given order, 4 fields, v0=(30,0,0), c0=340,
    absorbing, a Gaussian pulse exp(-r^2/0.01) at the origin (kernel comparisons across orders)."""
    m = dg.box_mesh(3, n, order, jitter=0.15, seed=seed)
    bc = dg.resolve_bc_sides(m, [(0x3F, dg.DG_BC_ABSORBING)])
    X = m.node_coords.reshape(-1, 3)
    u = np.zeros((4, len(X)))
    u[0] = np.exp(-(X ** 2).sum(axis=1) / 0.01)
    return dict(mesh=m, bc=bc, order=order, v0=(30.0, 0.0, 0.0), c0=340.0, fields=dg.DG_FIELDS_REF4, u0=u,
                describe=f"3D APE, {n}^3x6 jittered Kuhn tets (K={m.K}), p={order}, absorbing, Gaussian pulse")


WORKLOADS = {"dense3d": dense3d, "lee2d": lee2d, "box3d": box3d}


def run_closure(dim, steps, warmup=10, seed=5):
    n, p = (64, 3) if dim == 3 else (100, 4)
    m = dg.box_mesh(dim, n, p, jitter=0.15, seed=seed)
    X = m.node_coords.reshape(-1, 3)
    shape = f"{n}^3x6 jittered Kuhn tets" if dim == 3 else f"{n}^2x2 jittered triangles"
    out = dict(workload=f"closure{dim}d", steps=steps,
               describe=f"{dim}D APE, {shape} (K={m.K}), p={p}, mixed precision, boundary faces absorbing or learned")
    nfp = dg.num_nodes(dim - 1, p) if dim == 3 else p + 1
    params = dg.to_bf16(dg.closure_init(nfp, seed=seed, out_scale=0.1, dim=dim))
    zero = dg.to_bf16(dg.closure_init(nfp, seed=seed, out_scale=0.0, dim=dim))
    sides = 0x3F if dim == 3 else 0xF
    for mode, code, net in (("characteristic", dg.DG_BC_ABSORBING, None), ("learned_zero", dg.DG_BC_LEARNED, zero),
                            ("learned", dg.DG_BC_LEARNED, params)):
        bc = dg.resolve_bc_sides(m, [(sides, code)])
        s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=340.0,
                      precision=dg.DG_MIXED, steps_per_graph=20)
        if net is not None:
            s.set_closure(net)
        s.add_monopole(0.0, 0.0, 0.0, 0.05, 0.1, 200.0, 0.0, 1e9)
        for k in range(8):  # Fibonacci points on r = 0.8 (3D) / equally spaced on the circle (2D)
            if dim == 3:
                z = 1.0 - (2 * k + 1) / 8.0
                r, ph = np.sqrt(1.0 - z * z), np.pi * (3.0 - np.sqrt(5.0)) * k
                s.add_observer(0.8 * r * np.cos(ph), 0.8 * r * np.sin(ph), 0.8 * z, 0.03)
            else:
                s.add_observer(0.8 * np.cos(np.pi * k / 4), 0.8 * np.sin(np.pi * k / 4), 0.0, 0.03)
        info = s.info()
        dt = info.min_inradius / ((340.0 + 30.0) * (2 * p + 1))
        s.set_state(np.zeros((4, len(X))))
        s.run_steps(warmup, dt)
        s.set_state(np.zeros((4, len(X))))
        r = s.run_steps(steps, dt)
        stage, _, post = s.profile_steps(dt, 5)
        dofs = info.num_elements * info.nodes_per_element * 4
        out[mode] = dict(dof_updates_per_s=dofs * steps / (r.device_ms * 1e-3), ms_per_step=r.device_ms / steps,
                         stage_ms=[float(v) for v in stage], post_ms=float(post), dt=dt, dofs=int(dofs),
                         finite=bool(np.isfinite(r.observers).all()))
        s.close()
        print(f"# {mode}: {out[mode]['dof_updates_per_s']:.4e} DOF-updates/s, {out[mode]['ms_per_step']:.4f} ms/step",
              flush=True)
    out["closure_ms_per_step"] = out["learned"]["ms_per_step"] - out["characteristic"]["ms_per_step"]
    return out


def run_closure_train(samples=4_000_000, epochs=3, batch=8192, nfp=10, dim=3, seed=7):
    """Closure training throughput: Adam epochs over synthetic samples of the 3D order-3 face
    network (32 inputs, two hidden layers of 64, 10 outputs) with each training kernel."""
    nin, nout = dg.closure_sizes(nfp, dim)
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (samples, nin)).astype(np.float32)
    Y = (0.5 * np.tanh(X @ rng.normal(0, 0.3, (nin, nout)))).astype(np.float32)
    out = dict(workload="closure_train", samples=samples, epochs=epochs, batch=batch, nin=nin, nout=nout,
               describe=f"Adam on {samples} synthetic samples, network {nin}-64-64-{nout}, batch {batch}")
    for kernel in ("mfma", "valu"):
        os.environ["DGFEM_CLOSURE_TRAIN"] = kernel
        m0 = dg.closure_init(nfp, seed=1, out_scale=1.0, dim=dim)
        dg.closure_train(m0, X[:batch * 4], Y[:batch * 4], epochs=1, batch=batch, lr=1e-3, seed=2)  # warm-up
        t0 = time.perf_counter()
        _, _, hist = dg.closure_train(m0, X, Y, epochs=epochs, batch=batch, lr=1e-3, seed=2)
        sec = time.perf_counter() - t0
        out[kernel] = dict(seconds=sec, samples_per_s=samples * epochs / sec, loss=[float(v) for v in hist])
        print(f"# {kernel}: {samples * epochs / sec:.3e} samples/s ({sec:.2f} s), loss {hist[0]:.4e} -> {hist[-1]:.4e}",
              flush=True)
    os.environ.pop("DGFEM_CLOSURE_TRAIN", None)
    return out


def run(name, steps, warmup=10, mixed_kernels=("valu",), fp64=True, fp64_kernel="valu", **kw):
    """FP64 (unless fp64 is False; stage kernel fp64_kernel) and mixed precision with each
    stage kernel of mixed_kernels (DGFEM_KERNEL values), from the same initial state; errors
    against FP64."""
    w = WORKLOADS[name](**kw)
    m = w["mesh"]
    out = dict(workload=name, describe=w["describe"], steps=steps)
    states = {}
    runs = ([("fp64", dg.DG_FP64, fp64_kernel)] if fp64 else []) + [("mixed_" + k, dg.DG_MIXED, k) for k in mixed_kernels]
    for label, prec, kernel in runs:
        os.environ["DGFEM_KERNEL"] = kernel
        s = dg.Solver(m.node_coords, m.element_vertices, w["order"], w["bc"], v0=w["v0"], rho0=RHO0, c0=w["c0"],
                      fields=w["fields"], precision=prec, steps_per_graph=20)
        for ob in w.get("observers", ()):
            s.add_observer(*ob)
        info = s.info()
        dt = info.min_inradius / ((w["c0"] + np.linalg.norm(w["v0"])) * (2 * w["order"] + 1))
        s.set_state(w["u0"])
        s.run_steps(warmup, dt)           # graph capture and warm-up
        s.set_state(w["u0"])
        t0 = time.perf_counter()
        r = s.run_steps(steps, dt)
        wall = time.perf_counter() - t0
        dofs = info.num_elements * info.nodes_per_element * s.nf
        u = s.get_state()
        states[label] = (u[0].copy(), r.observers[:, :, 0].copy() if r.observers.size else None)
        del u
        stage, _, post = s.profile_steps(dt, 5)  # advances the state (taken above)
        out[label] = dict(dof_updates_per_s=dofs * steps / (r.device_ms * 1e-3), ms_per_step=r.device_ms / steps,
                          wall_ms_per_step=1e3 * wall / steps, stage_ms=[float(v) for v in stage],
                          post_ms=float(post), dt=dt, dofs=int(dofs), device_bytes=int(info.device_bytes),
                          kernel=dg.KERNEL_NAMES[info.stage_kernel])
        s.close()
        print(f"# {label}: {out[label]['dof_updates_per_s']:.4e} DOF-updates/s, {out[label]['ms_per_step']:.4f} "
              f"ms/step", flush=True)
    os.environ.pop("DGFEM_KERNEL", None)
    if fp64:
        up, _ = states["fp64"]
        for label, _, _ in runs[1:]:
            um, _ = states[label]
            out[label]["rel_l2_pressure_vs_fp64"] = float(np.linalg.norm(um - up) / np.linalg.norm(up))
            om, op = states[label][1], states["fp64"][1]
            if op is not None:
                out[label]["rel_l2_observer_vs_fp64"] = float(np.linalg.norm(om - op) / np.linalg.norm(op))
            out[label]["speedup_vs_fp64"] = out[label]["dof_updates_per_s"] / out["fp64"]["dof_updates_per_s"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["closure2d", "closure3d", "closure_train"],
                    default="dense3d")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--mixed-kernels", default="valu", help="comma-separated DGFEM_KERNEL values for the mixed runs")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--fp64-kernel", default="valu", help="DGFEM_KERNEL value of the FP64 run")
    ap.add_argument("--n", type=int, default=None, help="box3d: hexes per side")
    ap.add_argument("--order", type=int, default=None, help="box3d: Lagrange order")
    a = ap.parse_args(argv)
    if a.workload == "closure_train":
        res = run_closure_train()
    elif a.workload.startswith("closure"):
        res = run_closure(int(a.workload[7]), a.steps)
    else:
        kw = {k: v for k, v in (("n", a.n), ("order", a.order)) if v is not None}
        res = run(a.workload, a.steps, mixed_kernels=tuple(a.mixed_kernels.split(",")), fp64=not a.no_fp64,
                  fp64_kernel=a.fp64_kernel, **kw)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
