"""Learned absorbing-boundary closure: training data from FP64 runs on an enlarged domain,
training with the library's GPU kernels, and evaluation against the characteristic
absorbing BC of the reference (src/Mesh.cpp:619-666, :954-969) on the truncated domain.

The truncated mesh is the part of the enlarged mesh inside [-1,1]^dim, element for element
(vertices on the lines / planes x = +-1, y = +-1 (, z = +-1) are kept on them), so that a
truncated run differs from the enlarged one only by what happens on the truncation faces.
2D: triangles, monopoles or Gaussian pulses, observers on a circle. 3D: Kuhn tetrahedra,
dipole or quadrupole sources (expanded into antiphase monopoles as the reference's config
parser does, src/configParser.cpp:87-108), observers at Fibonacci points of a sphere. The network, its
training kernels and the in-solver inference live in libdgfem_amd.so (dg_closure_*,
dg_set_closure); this module drives them.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

import dgfem_amd as dg


@dataclass
class ClosureCase:
    n_trunc: int = 100          # cells per side of [-1,1]^dim (h = 2 / n_trunc)
    factor: int = 3             # enlarged domain [-factor, factor]^dim
    order: int = 4
    dim: int = 2
    jitter: float = 0.2
    seed: int = 4
    c0: float = 340.0
    rho0: float = 1.225
    source_size: float = 0.05   # monopole node radius (3D: dipole / quadrupole pole offset)
    n_observers: int = 8        # on the circle / sphere r = observer_radius
    observer_radius: float = 0.8
    observer_size: float = 0.03
    cfl: float = 1.0            # dt = cfl * r_min / ((c0 + |v0|max) (2p + 1))
    vmax: float = 0.3           # |v0| <= vmax * c0
    precision: int = dg.DG_FP64  # precision of the truncated-domain runs
    travel: float = 3.0         # simulated time = travel / c0 (source -> boundary -> observers)
    steps: int = 0              # 0: from travel and dt


@dataclass
class Meshes:
    parent_coords: np.ndarray
    parent_verts: np.ndarray
    sub_coords: np.ndarray
    sub_verts: np.ndarray
    sub_to_parent: np.ndarray
    faces_sub: tuple            # (element, local face) of the truncation faces in the truncated mesh
    faces_parent: tuple         # the same faces in the enlarged mesh
    r_min: float = 0.0


def _snap(mesh: dg.BoxMesh, lo: float, h: float, lines=(-1.0, 1.0)) -> np.ndarray:
    """Node coordinates with every vertex whose grid line (plane) is x = +-1, y = +-1 (or
    z = +-1) put back on it; the other nodes are rebuilt from the element's (moved) vertices."""
    X = mesh.node_coords
    K, Np, _ = X.shape
    d = mesh.dim
    V = X[:, :d + 1, :d]
    # barycentric coordinates of the element nodes (same local pattern for every element)
    A = np.stack([V[0, k + 1] - V[0, 0] for k in range(d)], axis=1)
    xi = np.linalg.solve(A, (X[0, :, :d] - V[0, 0]).T).T
    lam = np.concatenate([1 - xi.sum(axis=1, keepdims=True), xi], axis=1)
    rebuilt = np.einsum("ak,ekc->eac", lam, V)
    assert np.abs(rebuilt - X[:, :, :d]).max() < 1e-10, "generator node pattern is not affine-uniform"
    Vs = V.copy()
    for c in range(d):
        g = np.rint((Vs[..., c] - lo) / h)
        for ln in lines:
            on = np.abs(g * h + lo - ln) < 1e-9
            Vs[..., c][on] = ln
    out = X.copy()
    out[:, :, :d] = np.einsum("ak,ekc->eac", lam, Vs)
    return out


def _inradius(V):
    """Inradius of triangles [K][3][2] or tetrahedra [K][4][3]."""
    if V.shape[1] == 3:
        a = np.linalg.norm(V[:, 1] - V[:, 2], axis=1)
        b = np.linalg.norm(V[:, 0] - V[:, 2], axis=1)
        c = np.linalg.norm(V[:, 0] - V[:, 1], axis=1)
        area = 0.5 * np.abs(_cross2(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]))
        return 2 * area / (a + b + c)
    vol = np.abs(np.einsum("ij,ij->i", V[:, 1] - V[:, 0], np.cross(V[:, 2] - V[:, 0], V[:, 3] - V[:, 0]))) / 6

    def tri(p, q, r):
        return 0.5 * np.linalg.norm(np.cross(q - p, r - p), axis=1)
    S = tri(V[:, 0], V[:, 1], V[:, 2]) + tri(V[:, 0], V[:, 1], V[:, 3]) + tri(V[:, 0], V[:, 2], V[:, 3]) + \
        tri(V[:, 1], V[:, 2], V[:, 3])
    return 3 * vol / S


def _cross2(a, b):
    return a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]


def build_meshes(case: ClosureCase) -> Meshes:
    F = case.factor
    d = case.dim
    n = case.n_trunc * F
    h = 2.0 / case.n_trunc
    m = dg.box_mesh(d, n, case.order, jitter=case.jitter, seed=case.seed, lo=-float(F), hi=float(F))
    coords = _snap(m, -float(F), h)
    cen = coords[:, :d + 1, :d].mean(axis=1)
    inside = np.where(np.all(np.abs(cen) < 1, axis=1))[0]
    sub_coords = np.ascontiguousarray(coords[inside])
    sub_verts = np.ascontiguousarray(m.element_vertices[inside])
    pre = dg.mesh_precompute(sub_coords, sub_verts, case.order)
    e, f = np.nonzero(pre["nbr"] < 0)
    r_min = float(min(_inradius(sub_coords[:, :d + 1, :d]).min(), _inradius(coords[:, :d + 1, :d]).min()))
    return Meshes(coords, m.element_vertices, sub_coords, sub_verts, inside, (e.astype(np.int64), f.astype(np.int32)),
                  (inside[e].astype(np.int64), f.astype(np.int32)), r_min)


def observers(case: ClosureCase):
    r = case.observer_radius
    if case.dim == 2:
        ang = 2 * np.pi * (np.arange(case.n_observers) + 0.5) / case.n_observers
        return [(r * np.cos(a), r * np.sin(a), 0.0, case.observer_size) for a in ang]
    # Fibonacci points of the sphere
    k = np.arange(case.n_observers) + 0.5
    z = 1 - 2 * k / case.n_observers
    phi = np.pi * (3 - np.sqrt(5)) * k
    rho = np.sqrt(1 - z * z)
    return [(r * rho[i] * np.cos(phi[i]), r * rho[i] * np.sin(phi[i]), r * z[i], case.observer_size)
            for i in range(case.n_observers)]


def time_step(case: ClosureCase, meshes: Meshes) -> tuple[float, int]:
    dt = case.cfl * meshes.r_min / ((1 + case.vmax) * case.c0 * (2 * case.order + 1))
    steps = case.steps or int(np.ceil(case.travel / case.c0 / dt))
    return dt, steps


def scenario(case: ClosureCase, seed: int) -> dict:
    """A monopole in the disk r < 0.5 at 20-200 Hz, or a Gaussian pulse of width 0.05-0.2,
    with a mean flow of |v0| <= vmax c0 in a random direction."""
    rng = np.random.default_rng(seed)
    if case.dim == 3:
        return _scenario3d(case, rng)
    r, a = 0.5 * np.sqrt(rng.uniform()), rng.uniform(0, 2 * np.pi)
    va, vm = rng.uniform(0, 2 * np.pi), case.vmax * case.c0 * np.sqrt(rng.uniform())
    s = dict(x=r * np.cos(a), y=r * np.sin(a), v0=(vm * np.cos(va), vm * np.sin(va), 0.0))
    if rng.uniform() < 0.5:
        s.update(kind="monopole", freq=rng.uniform(20, 200), amp=1.0)
    else:
        s.update(kind="gaussian", width=rng.uniform(0.05, 0.2), amp=1.0)
    return s


def _scenario3d(case: ClosureCase, rng) -> dict:
    """3D: a dipole or a quadrupole at 50-300 Hz centred in the ball r < 0.5, mean flow
    |v0| <= vmax c0 in a random direction."""
    c = rng.normal(0, 1, 3)
    c *= 0.5 * rng.uniform() ** (1 / 3) / np.linalg.norm(c)
    d = rng.normal(0, 1, 3)
    vm = case.vmax * case.c0 * rng.uniform() ** (1 / 3)
    return dict(kind="dipole" if rng.uniform() < 0.5 else "quadrupole", x=c[0], y=c[1], z=c[2],
                freq=rng.uniform(50, 300), amp=1.0, v0=tuple(vm * d / np.linalg.norm(d)))


def poles(scen, size):
    """Monopoles (x, y, z, radius, phase) of a dipole / quadrupole of pole offset `size`, as the
    reference's parser expands them (src/configParser.cpp:87-108): a dipole is two monopoles of
    radius size/2 at x -+ size in antiphase; a quadrupole adds the pair at y -+ size in antiphase
    with the first."""
    x, y, z = scen["x"], scen["y"], scen["z"]
    out = [(x - size, y, z, size / 2, 0.0), (x + size, y, z, size / 2, np.pi)]
    if scen["kind"] == "quadrupole":
        out = [(x - size, y, z, size / 2, 0.0), (x + size, y, z, size / 2, 0.0),
               (x, y - size, z, size / 2, np.pi), (x, y + size, z, size / 2, np.pi)]
    return out


def _solver(coords, verts, case, scen, face_bc=None, precision=dg.DG_FP64):
    s = dg.Solver(coords, verts, case.order, face_bc, v0=scen["v0"], rho0=case.rho0, c0=case.c0, precision=precision,
                  steps_per_graph=16)
    if scen["kind"] == "monopole":
        s.add_monopole(scen["x"], scen["y"], 0.0, case.source_size, scen["amp"], scen["freq"], 0.0, 1e9)
    elif scen["kind"] in ("dipole", "quadrupole"):
        for (px, py, pz, rad, ph) in poles(scen, case.source_size):
            s.add_monopole(px, py, pz, rad, scen["amp"], scen["freq"], ph, 1e9)
    else:
        s.add_gaussian(scen["x"], scen["y"], 0.0, scen["width"] ** 2, scen["amp"])
    for o in observers(case):
        s.add_observer(*o)
    return s


def enlarged_run(case, meshes, scen, dt, steps, record_every=1, face_fraction=1.0, seed=0, min_scale=0.0,
                 target_clip=10.0, log=None, progress_every=0):
    """FP64 run on the enlarged mesh. Returns the observer pressure series [steps][nobs] and
    training samples (X, Y) from the truncation-face traces every record_every steps (faces
    whose trace scale is below min_scale times the run's largest are dropped). With log and
    progress_every, a line every progress_every steps."""
    s = _solver(meshes.parent_coords, meshes.parent_verts, case, scen)
    e, f = meshes.faces_parent
    rng = np.random.default_rng(seed)
    Xs, Ys, Ss, obs = [], [], [], []
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        if record_every > 0:
            k = min(record_every, steps - done)
        else:
            k = min(progress_every, steps - done) if (log and progress_every > 0) else steps
        if record_every > 0:
            sel = np.arange(e.size) if face_fraction >= 1 else np.sort(
                rng.choice(e.size, max(1, int(face_fraction * e.size)), replace=False))
            tr, nrm = s.face_traces(e[sel], f[sel])
            X, Y, S = dg.closure_features(tr, nrm, scen["v0"], case.rho0, case.c0, dim=case.dim)
            Xs.append(X)
            Ys.append(Y)
            Ss.append(S)
        r = s.run_steps(k, dt, t_start=done * dt, residuals=False)
        obs.append(r.observers[:, :, 0])
        if log and progress_every > 0 and (done + k) // progress_every > done // progress_every:
            log(f"    enlarged run: {done + k}/{steps} steps ({time.perf_counter() - t0:.1f} s)")
        done += k
    s.close()
    if not Xs:
        return np.concatenate(obs), None, None
    X, Y, S = np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ss)
    # optionally drop faces the wave has not reached; targets of near-zero traces are clipped
    keep = S > min_scale * S.max()
    Y = np.clip(Y[keep], -target_clip, target_clip)
    return np.concatenate(obs), X[keep], Y


def truncated_run(case, meshes, scen, dt, steps, mode, params=None, precision=None):
    """mode: 'characteristic' (the reference's absorbing BC) or 'learned' (params: BF16
    network; None = the zero network, i.e. the exact characteristic condition)."""
    K = meshes.sub_verts.shape[0]
    precision = case.precision if precision is None else precision
    bc = np.zeros((K, case.dim + 1), dtype=np.int8)
    if mode == "learned":
        e, f = meshes.faces_sub
        bc[e, f] = dg.DG_BC_LEARNED
    s = _solver(meshes.sub_coords, meshes.sub_verts, case, scen, bc, precision)
    if mode == "learned" and params is not None:
        s.set_closure(params)
    r = s.run_steps(steps, dt, residuals=False)
    ms = r.device_ms
    s.close()
    return r.observers[:, :, 0], ms


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@dataclass
class Report:
    case: dict
    dt: float
    steps: int
    train_samples: int
    epoch_loss: list
    train_seconds: float
    alpha: float = 1.0
    validation: dict = field(default_factory=dict)
    eval: list = field(default_factory=list)


def scale_output(master: np.ndarray, nfp: int, alpha: float, dim: int = 2) -> np.ndarray:
    """Network with its output layer (W3, b3) multiplied by alpha: the learned incoming
    characteristic is blended with the exact characteristic condition (alpha = 0)."""
    nin, _ = dg.closure_sizes(nfp, dim)
    off = 64 * nin + 64 + 64 * 64 + 64
    m = np.array(master, dtype=np.float32, copy=True)
    m[off:] *= alpha
    return m


def collect_training(case: ClosureCase, meshes: Meshes, dt: float, steps: int, n_train: int, record_every: int = 2,
                     face_fraction: float = 1.0, log=print, progress_every: int = 0, min_scale: float = 0.0):
    """FP64 enlarged-domain training runs (seeds 1000+i) -> samples (X, Y); samples whose trace
    scale is below min_scale times the run's largest (faces the wave has not reached) are dropped."""
    Xs, Ys = [], []
    for i in range(n_train):
        scen = scenario(case, 1000 + i)
        t0 = time.perf_counter()
        _, X, Y = enlarged_run(case, meshes, scen, dt, steps, record_every, face_fraction, seed=i, log=log,
                               progress_every=progress_every, min_scale=min_scale)
        Xs.append(X)
        Ys.append(Y)
        log(f"  training run {i + 1}/{n_train} ({scen['kind']}): {X.shape[0]} samples "
            f"({time.perf_counter() - t0:.1f} s)")
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    return X, Y


def train_network(case: ClosureCase, X, Y, epochs: int, batch: int, lr: float, log=print):
    """Adam on the GPU from the seeded initial network; returns (master, loss history, seconds)."""
    master = dg.closure_init(face_nodes(case), seed=case.seed, dim=case.dim)
    t0 = time.perf_counter()
    master, _, hist = dg.closure_train(master, X, Y, epochs=epochs, batch=batch, lr=lr, seed=case.seed)
    t_train = time.perf_counter() - t0
    log(f"  trained {epochs} epochs on {X.shape[0]} samples in {t_train:.1f} s: loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    return master, [float(v) for v in hist], t_train


def face_nodes(case: ClosureCase) -> int:
    return case.order + 1 if case.dim == 2 else (case.order + 1) * (case.order + 2) // 2


def choose_alpha(case: ClosureCase, meshes: Meshes, dt: float, steps: int, master, n_val: int,
                 alphas=(0.0, 0.25, 0.5, 0.75, 1.0), log=print, progress_every: int = 0):
    """Output blend chosen on validation runs (seeds 3000+j) by the mean observer error."""
    val = []
    for j in range(n_val):
        scen = scenario(case, 3000 + j)
        ref, _, _ = enlarged_run(case, meshes, scen, dt, steps, record_every=0, log=log, progress_every=progress_every)
        val.append((scen, ref))
    best, table = None, {}
    for a in alphas:
        w = dg.to_bf16(scale_output(master, face_nodes(case), a, case.dim))
        errs = [rel_l2(truncated_run(case, meshes, sc, dt, steps, "learned", w)[0], ref) for sc, ref in val]
        table[str(a)] = errs
        log(f"  validation alpha={a}: observer rel L2 {['%.3e' % e for e in errs]}")
        if best is None or np.mean(errs) < best[0]:
            best = (float(np.mean(errs)), a)
    return best[1], table


def evaluate(case: ClosureCase, meshes: Meshes, dt: float, steps: int, params, seeds, log=print,
             progress_every: int = 0) -> list:
    """Held-out runs: the reference's characteristic BC, the exact characteristic condition
    (zero network) and the learned closure on the truncated domain, each against the
    enlarged-domain FP64 solve (observer pressure rel-L2). One dict per seed."""
    out_all = []
    for j, seed in enumerate(seeds):
        scen = scenario(case, seed)
        t0 = time.perf_counter()
        ref, _, _ = enlarged_run(case, meshes, scen, dt, steps, record_every=0, log=log, progress_every=progress_every)
        t_ref = time.perf_counter() - t0
        out = dict(seed=int(seed), kind=scen["kind"], v0=[float(v) for v in scen["v0"]], enlarged_seconds=t_ref)
        for name, mode, prm in (("characteristic", "characteristic", None),
                                ("exact_characteristic", "learned", None), ("learned", "learned", params)):
            o, ms = truncated_run(case, meshes, scen, dt, steps, mode, prm)
            out[name] = rel_l2(o, ref)
            out[name + "_ms_per_step"] = ms / steps
        out_all.append(out)
        log(f"  held-out {j + 1}/{len(seeds)} (seed {seed}, {scen['kind']}): observer rel L2 vs enlarged: "
            f"characteristic {out['characteristic']:.3e} exact-characteristic {out['exact_characteristic']:.3e} "
            f"learned {out['learned']:.3e}")
    return out_all


def train_and_evaluate(case: ClosureCase, n_train: int, n_eval: int, n_val: int = 2, epochs: int = 60,
                       batch: int = 2048, lr: float = 1e-3, record_every: int = 2, face_fraction: float = 1.0,
                       alphas=(0.0, 0.25, 0.5, 0.75, 1.0), log=print) -> Report:
    """Training runs (seeds 1000+i) -> samples -> Adam on the GPU; the output blend alpha is
    chosen on validation runs (seeds 3000+j) by the mean observer error; held-out runs
    (seeds 5000+j) compare the reference's characteristic BC, the exact characteristic
    condition (zero network) and the learned closure against the enlarged-domain solve."""
    meshes = build_meshes(case)
    dt, steps = time_step(case, meshes)
    log(f"closure case: enlarged K={meshes.parent_verts.shape[0]} truncated K={meshes.sub_verts.shape[0]} "
        f"truncation faces={meshes.faces_sub[0].size} dt={dt:.3e} steps={steps}")
    X, Y = collect_training(case, meshes, dt, steps, n_train, record_every, face_fraction, log)
    master, hist, t_train = train_network(case, X, Y, epochs, batch, lr, log)
    rep = Report(case=case.__dict__.copy(), dt=dt, steps=steps, train_samples=int(X.shape[0]),
                 epoch_loss=hist, train_seconds=t_train)
    del X, Y
    rep.alpha, rep.validation = choose_alpha(case, meshes, dt, steps, master, n_val, alphas, log)
    params = dg.to_bf16(scale_output(master, face_nodes(case), rep.alpha, case.dim))
    rep.params = params
    rep.master = master
    rep.eval = evaluate(case, meshes, dt, steps, params, [5000 + j for j in range(n_eval)], log)
    return rep


def save_network(path: str, master, case: ClosureCase, alpha=None, **info):
    """FP32 master weights with the case they were trained on (a plain .npz, no pickled objects)."""
    import json
    np.savez(path, master=np.asarray(master, dtype=np.float32),
             alpha=np.float64(np.nan if alpha is None else alpha),
             meta=np.array(json.dumps(dict(case=case.__dict__, **info))))


def load_network(path: str):
    import json
    with np.load(path, allow_pickle=False) as z:
        alpha = float(z["alpha"])
        return np.array(z["master"], dtype=np.float32), (None if np.isnan(alpha) else alpha), json.loads(str(z["meta"]))


def main(argv=None):
    """Whole pipeline in one process, or split across processes: --train N --save-network F
    (training only with --eval 0 --val 0), then --load-network F [--alpha A] --eval-seeds ...
    (validation and held-out runs)."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="Train the learned absorbing closure on enlarged-domain runs and "
                                             "compare it with the characteristic BC on the truncated domain.")
    ap.add_argument("--dim", type=int, default=2, choices=(2, 3))
    ap.add_argument("--n-trunc", type=int, default=100, help="cells per side of [-1,1]^dim (h = 2/n)")
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--factor", type=int, default=3, help="enlarged domain [-factor, factor]^dim")
    ap.add_argument("--jitter", type=float, default=0.2, help="interior vertex jitter, fraction of h")
    ap.add_argument("--steps", type=int, default=0, help="RK4 steps per run (0: 3 m of travel at c0)")
    ap.add_argument("--source-size", type=float, default=0.05,
                    help="monopole radius (2D) / dipole-quadrupole pole offset (3D)")
    ap.add_argument("--mixed", action="store_true", help="truncated-domain runs in mixed precision")
    ap.add_argument("--train", type=int, default=32)
    ap.add_argument("--val", type=int, default=2)
    ap.add_argument("--eval", type=int, default=4, help="held-out runs, seeds 5000, 5001, ...")
    ap.add_argument("--eval-seeds", default="", help="comma-separated held-out seeds (overrides --eval)")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--record-every", type=int, default=2)
    ap.add_argument("--face-fraction", type=float, default=1.0)
    ap.add_argument("--observer-size", type=float, default=0.03)
    ap.add_argument("--alpha", type=float, default=None, help="output blend (skips the validation runs)")
    ap.add_argument("--alphas", default="0,0.25,0.5,0.75,1", help="output blends tried on the validation runs")
    ap.add_argument("--save-network", default="", help="write the trained FP32 master weights (.npz)")
    ap.add_argument("--load-network", default="", help="skip training; use these weights (.npz)")
    ap.add_argument("--progress-every", type=int, default=1000, help="enlarged runs: a line every N steps")
    ap.add_argument("--min-scale", type=float, default=0.0,
                    help="drop training samples whose trace scale is below this fraction of the run's largest")
    ap.add_argument("--out", default="", help="write the report as JSON here")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    case = ClosureCase(n_trunc=a.n_trunc, order=a.order, factor=a.factor, observer_size=a.observer_size, dim=a.dim,
                       source_size=a.source_size, jitter=a.jitter, steps=a.steps,
                       precision=dg.DG_MIXED if a.mixed else dg.DG_FP64)
    seeds = [int(v) for v in a.eval_seeds.split(",") if v] or [5000 + j for j in range(a.eval)]
    split = bool(a.save_network or a.load_network or a.alpha is not None or a.eval_seeds)
    if not split:
        rep = train_and_evaluate(case, a.train, a.eval, a.val, a.epochs, a.batch, a.lr, a.record_every,
                                 a.face_fraction, log=log)
        out = dict(rep.__dict__)
        out.pop("params", None)
        out.pop("master", None)
    else:
        t0 = time.perf_counter()
        meshes = build_meshes(case)
        dt, steps = time_step(case, meshes)
        log(f"closure case: enlarged K={meshes.parent_verts.shape[0]} truncated K={meshes.sub_verts.shape[0]} "
            f"truncation faces={meshes.faces_sub[0].size} dt={dt:.3e} steps={steps} "
            f"(meshes built in {time.perf_counter() - t0:.1f} s)")
        out = dict(case=case.__dict__.copy(), dt=dt, steps=steps)
        alpha = a.alpha
        if a.load_network:
            master, saved_alpha, meta = load_network(a.load_network)
            alpha = saved_alpha if alpha is None else alpha
            out["network"] = meta
        else:
            X, Y = collect_training(case, meshes, dt, steps, a.train, a.record_every, a.face_fraction, log,
                                    a.progress_every, a.min_scale)
            master, hist, t_train = train_network(case, X, Y, a.epochs, a.batch, a.lr, log)
            out.update(train_samples=int(X.shape[0]), epoch_loss=hist, train_seconds=t_train)
            del X, Y
        if alpha is None and a.val > 0:
            alpha, out["validation"] = choose_alpha(case, meshes, dt, steps, master, a.val,
                                                    tuple(float(v) for v in a.alphas.split(",")), log=log,
                                                    progress_every=a.progress_every)
        if a.save_network:
            save_network(a.save_network, master, case, alpha, train_runs=a.train, epochs=a.epochs,
                         record_every=a.record_every, face_fraction=a.face_fraction, min_scale=a.min_scale)
            log(f"  network written to {a.save_network} (alpha {alpha})")
        out["alpha"] = alpha
        if seeds and a.eval > 0 or a.eval_seeds:
            if alpha is None:
                raise SystemExit("held-out runs need an output blend: --alpha, --val > 0 or a saved one")
            params = dg.to_bf16(scale_output(master, face_nodes(case), alpha, case.dim))
            out["eval"] = evaluate(case, meshes, dt, steps, params, seeds, log, a.progress_every)
    if out.get("eval"):
        for k in ("characteristic", "exact_characteristic", "learned"):
            out["mean_" + k] = float(np.mean([e[k] for e in out["eval"]]))
    print(json.dumps({k: out[k] for k in out if k.startswith("mean_") or k in ("alpha", "steps", "dt",
                                                                              "train_samples")}), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
