"""GPU tests of the learned absorbing-boundary closure: the library's BF16 forward/backward
kernels against the FP64 oracle (oracle/closure_oracle.py), training, the in-solver
inference on DG_BC_LEARNED faces against the NumPy nodal restatement with the oracle's
exterior states, and the gain over the characteristic BC on a reduced closure case.

Bars: forward 2e-5 relative (FP32 accumulation of BF16 weights vs FP64 with the same
weights); loss 1e-5 relative, gradient 1e-4 relative L2; RHS with learned faces 1e-6
relative L2 per field (the exterior states carry the FP32 network error).
"""
import numpy as np
import pytest

import closure_oracle as co
import dgfem_amd as dg
import nodal_ref
from cases import rel_l2, smooth_state

pytestmark = pytest.mark.gpu


def _params(nfp, seed, scale=0.3, dim=2):
    rng = np.random.default_rng(seed)
    m = rng.normal(0, scale, dg.closure_num_params(nfp, dim)).astype(np.float32)
    bits = dg.to_bf16(m)
    return bits, co.bf16_to_f64(bits)


@pytest.mark.parametrize("nfp,dim", [(2, 2), (4, 2), (5, 2), (6, 2), (3, 3), (6, 3), (10, 3)])
def test_forward_matches_oracle(nfp, dim):
    print(f"case closure forward nfp={nfp} dim={dim}", flush=True)
    nin, nout = co.sizes(nfp, dim)
    bits, p64 = _params(nfp, nfp, dim=dim)
    X = np.random.default_rng(1).normal(0, 1, (3000, nin)).astype(np.float32)
    y = dg.closure_forward(bits, X)
    yo, _ = co.forward(p64, X.astype(np.float64))
    assert np.abs(y - yo).max() <= 2e-5 * max(1.0, np.abs(yo).max())


@pytest.mark.parametrize("train_kernel", ["mfma", "valu"])
@pytest.mark.parametrize("nfp,n,dim", [(3, 1, 2), (3, 777, 2), (5, 5000, 2), (6, 64, 3), (10, 3001, 3)])
def test_loss_and_gradient_match_oracle(nfp, n, dim, train_kernel, monkeypatch):
    """Both training kernels: the matrix-core one (split-BF16 products, the default) and the
    one-sample-per-wave VALU one; sample counts that fill no, one or many 32-sample tiles."""
    print(f"case closure grad nfp={nfp} n={n} dim={dim} kernel={train_kernel}", flush=True)
    monkeypatch.setenv("DGFEM_CLOSURE_TRAIN", train_kernel)
    nin, nout = co.sizes(nfp, dim)
    bits, p64 = _params(nfp, 10 + nfp, dim=dim)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (n, nin)).astype(np.float32)
    Y = rng.normal(0, 0.5, (n, nout)).astype(np.float32)
    loss, g = dg.closure_loss_grad(bits, X, Y)
    lo, go = co.loss_grad(p64, X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * lo
    assert rel_l2(g, go) < 1e-4, rel_l2(g, go)


@pytest.mark.parametrize("train_kernel", ["mfma", "valu"])
def test_training_reduces_loss_on_a_smooth_target(train_kernel, monkeypatch):
    print(f"case closure train synthetic kernel={train_kernel}", flush=True)
    monkeypatch.setenv("DGFEM_CLOSURE_TRAIN", train_kernel)
    nfp = 5
    nin, nout = co.sizes(nfp)
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (20000, nin)).astype(np.float32)
    A = rng.normal(0, 0.3, (nin, nout))
    Y = (0.5 * np.tanh(X @ A)).astype(np.float32)
    m0 = dg.closure_init(nfp, seed=1, out_scale=1.0)
    l0, _ = dg.closure_loss_grad(dg.to_bf16(m0), X, Y)
    m, w, hist = dg.closure_train(m0, X, Y, epochs=15, batch=1024, lr=3e-3, seed=2)
    l1, _ = dg.closure_loss_grad(w, X, Y)
    print(f"loss {l0:.3e} -> {l1:.3e}; epochs {hist}")
    assert np.all(np.isfinite(hist)) and l1 < 0.1 * l0
    assert np.array_equal(w, dg.to_bf16(m))  # BF16 copy of the FP32 master weights


def _learned_case(n, p, seed=5):
    m = dg.box_mesh(2, n, p, jitter=0.2, seed=seed)
    bc = np.where(m.face_side >= 0, dg.DG_BC_LEARNED, dg.DG_BC_ABSORBING).astype(np.int8)
    return m, bc


@pytest.mark.parametrize("p,precision", [(2, dg.DG_FP64), (4, dg.DG_FP64), (4, dg.DG_MIXED)])
def test_rhs_with_learned_faces_matches_nodal_restatement(p, precision):
    print(f"case closure rhs p={p} precision={precision}", flush=True)
    v0, rho0, c0 = (40.0, -25.0, 0.0), 1.225, 100.0
    m, bc = _learned_case(5, p)
    bits, p64 = _params(p + 1, 7, scale=0.2)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=rho0, c0=c0, precision=precision)
    e, f = s.learned_faces()
    assert e.size == 4 * 5
    s.set_closure(bits)
    u = smooth_state(m)
    k = s.rhs(u, 1e-4)
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, p, bc)
    ops = dg.ref_operators(2, p)
    kr = nodal_ref.nodal_rhs(u, pre, ops, 2, v0, rho0, c0, 1e-4,
                             closure=lambda qm, nn: co.exterior(qm, nn, p64, v0, rho0, c0))
    k0 = nodal_ref.nodal_rhs(u, pre, ops, 2, v0, rho0, c0, 1e-4,
                             closure=lambda qm, nn: co.exterior_from_incoming(qm, nn, 0.0, v0, rho0, c0))
    tol = 1e-6 if precision == dg.DG_FP64 else 2e-5
    for fld in range(3):
        assert rel_l2(k[fld], kr[fld]) < tol, (fld, rel_l2(k[fld], kr[fld]))
    # the network's contribution is resolved, not lost in the tolerance
    assert max(rel_l2(kr[fld], k0[fld]) for fld in range(3)) > 100 * tol


def test_zero_network_and_traces():
    print("case closure zero network, traces", flush=True)
    v0, rho0, c0 = (0.0, 30.0, 0.0), 1.225, 100.0
    p = 3
    m, bc = _learned_case(4, p)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=rho0, c0=c0)
    u = smooth_state(m)
    k = s.rhs(u, 1e-4)  # default network: all zero, i.e. the exact characteristic condition
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, p, bc)
    ops = dg.ref_operators(2, p)
    k0 = nodal_ref.nodal_rhs(u, pre, ops, 2, v0, rho0, c0, 1e-4,
                             closure=lambda qm, nn: co.exterior_from_incoming(qm, nn, 0.0, v0, rho0, c0))
    for fld in range(3):
        assert rel_l2(k[fld], k0[fld]) < 1e-12
    # traces of the state on interior and boundary faces, against the host-side gather
    s.set_state(u)
    K = m.K
    elem = np.repeat(np.arange(K), 3)
    face = np.tile(np.arange(3), K).astype(np.int32)
    tr, nrm = s.face_traces(elem, face)
    tc, fm, fperm = pre["to_canon"], ops["fmask"], pre["face_perm"]
    q = np.zeros((K, tc.shape[1], 4))
    q[np.arange(K)[:, None], tc, :] = u.T.reshape(K, tc.shape[1], 4)
    for i in range(0, elem.size, 7):
        e_, f_ = elem[i], face[i]
        np.testing.assert_array_equal(tr[i, 0], q[e_, fm[f_], :])
        nb = pre["nbr"][e_, f_]
        if nb >= 0:
            cd = pre["code"][e_, f_]
            idx = fm[cd & 3][fperm[cd >> 2]]
            np.testing.assert_array_equal(tr[i, 1], q[nb, idx, :])
        else:
            assert not tr[i, 1].any()
        assert abs(np.linalg.norm(nrm[i]) - 1) < 1e-12


def test_graph_replay_equals_eager_with_learned_faces():
    print("case closure graph vs eager", flush=True)
    v0, rho0, c0 = (20.0, 10.0, 0.0), 1.225, 100.0
    m, bc = _learned_case(6, 4)
    bits, _ = _params(5, 9, scale=0.2)
    out = []
    for graph in (True, False):
        s = dg.Solver(m.node_coords, m.element_vertices, 4, bc, v0=v0, rho0=rho0, c0=c0, use_graph=graph,
                      steps_per_graph=8)
        s.set_closure(bits)
        s.add_gaussian(0.1, -0.2, 0.0, 0.05, 1.0)
        s.add_observer(0.9, 0.0, 0.0, 0.1)
        r = s.run_steps(40, 2e-5)
        out.append((s.get_state(), r.observers))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_learned_closure_beats_characteristic_bc():
    """Reduced closure case: truncated [-1,1]^2 (20x20x2 triangles, p=4) inside the
    enlarged [-3,3]^2; 12 training runs, 2 validation runs (output blend), 3 held-out runs.
    Against the enlarged-domain observer signals the learned closure must be more accurate
    than the reference's characteristic BC on average, never blow up or fall far behind it on
    a single run, and not regress on average against the exact characteristic condition it
    starts from (at this size the learned gain over that condition is within run-to-run noise;
    the configured size is measured by dgfem-acoustic_amd/closure.py)."""
    import closure as cl
    case = cl.ClosureCase(n_trunc=20, observer_size=0.06)
    rep = cl.train_and_evaluate(case, n_train=12, n_eval=3, n_val=2, epochs=60, batch=2048, lr=1e-3,
                                record_every=2, log=lambda s: print(s, flush=True))
    print(f"alpha {rep.alpha}")
    assert rep.epoch_loss[-1] < rep.epoch_loss[0]
    learned = np.array([ev["learned"] for ev in rep.eval])
    ref_bc = np.array([ev["characteristic"] for ev in rep.eval])
    exact = np.array([ev["exact_characteristic"] for ev in rep.eval])
    assert np.all(np.isfinite(learned)) and np.all(learned < 1.25 * ref_bc), (learned, ref_bc)
    assert learned.mean() < ref_bc.mean(), (learned, ref_bc)
    assert learned.mean() <= 1.05 * exact.mean(), (learned, exact)


# ---- tetrahedra ----

def test_learned_closure_beats_characteristic_bc_3d():
    """Reduced 3D closure case: truncated [-1,1]^3 (8^3x6 jittered Kuhn tets, p=3, mixed
    precision) inside the enlarged [-2,2]^3 (FP64); dipole / quadrupole sources, 8 observers on
    the sphere r = 0.8; 12 training runs, 2 validation runs (output blend), 4 held-out runs.
    Against the enlarged-domain observer signals the learned closure must beat the reference's
    characteristic BC and the exact characteristic condition (the zero network) on average,
    and not fall far behind either on any single run."""
    import closure as cl
    case = cl.ClosureCase(dim=3, n_trunc=8, factor=2, order=3, jitter=0.15, source_size=0.25,
                          observer_size=0.12, precision=dg.DG_MIXED)
    rep = cl.train_and_evaluate(case, n_train=12, n_eval=4, n_val=2, epochs=40, batch=2048, lr=1e-3,
                                record_every=2, log=lambda s: print(s, flush=True))
    print(f"alpha {rep.alpha}")
    assert rep.epoch_loss[-1] < rep.epoch_loss[0]
    learned = np.array([ev["learned"] for ev in rep.eval])
    ref_bc = np.array([ev["characteristic"] for ev in rep.eval])
    exact = np.array([ev["exact_characteristic"] for ev in rep.eval])
    for a, b, c in zip(learned, exact, ref_bc):
        print(f"learned {a:.3e}  exact characteristic {b:.3e} ({a / b - 1:+.1%})  reference BC {c:.3e} "
              f"({a / c - 1:+.1%})")
    assert np.all(np.isfinite(learned)) and np.all(learned < 1.1 * exact) and np.all(learned < 1.1 * ref_bc)
    assert learned.mean() < exact.mean() and learned.mean() < ref_bc.mean(), (learned, exact, ref_bc)



def _learned_case3d(n, p, seed=6):
    m = dg.box_mesh(3, n, p, jitter=0.15, seed=seed)
    bc = np.where(m.face_side >= 0, dg.DG_BC_LEARNED, dg.DG_BC_ABSORBING).astype(np.int8)
    return m, bc


@pytest.mark.parametrize("p,precision,v0", [(1, dg.DG_FP64, (-10.0, 20.0, 5.0)), (2, dg.DG_FP64, (40.0, -25.0, 15.0)),
                                            (3, dg.DG_FP64, (0.0, 0.0, 0.0)), (3, dg.DG_MIXED, (30.0, 0.0, 0.0))])
def test_rhs_with_learned_faces_matches_nodal_restatement_3d(p, precision, v0):
    """Tetrahedra: the closure kernel's exterior states (3D face frame, two tangential
    components) against the oracle's, through the whole right-hand side. v0 = 0 exercises the
    axis-based tangents; (30, 0, 0) has faces with and without tangential mean flow."""
    print(f"case closure rhs 3d p={p} precision={precision} v0={v0}", flush=True)
    rho0, c0 = 1.225, 340.0
    m, bc = _learned_case3d(3, p)
    nfp = (p + 1) * (p + 2) // 2
    bits, p64 = _params(nfp, 8 + p, scale=0.2, dim=3)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=rho0, c0=c0, precision=precision)
    e, f = s.learned_faces()
    assert e.size == 6 * 9 * 2
    s.set_closure(bits)
    u = smooth_state(m)
    k = s.rhs(u, 1e-5)
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, p, bc)
    ops = dg.ref_operators(3, p)
    kr = nodal_ref.nodal_rhs(u, pre, ops, 3, v0, rho0, c0, 1e-5,
                             closure=lambda qm, nn: co.exterior(qm, nn, p64, v0, rho0, c0, dim=3))
    k0 = nodal_ref.nodal_rhs(u, pre, ops, 3, v0, rho0, c0, 1e-5,
                             closure=lambda qm, nn: co.exterior_from_incoming(qm, nn, 0.0, v0, rho0, c0, dim=3))
    tol = 1e-6 if precision == dg.DG_FP64 else 2e-5
    for fld in range(4):
        assert rel_l2(k[fld], kr[fld]) < tol, (fld, rel_l2(k[fld], kr[fld]))
    assert max(rel_l2(kr[fld], k0[fld]) for fld in range(4)) > 100 * tol
    # the default (zero) network is the exact characteristic condition
    s2 = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=rho0, c0=c0, precision=precision)
    k2 = s2.rhs(u, 1e-5)
    for fld in range(4):
        assert rel_l2(k2[fld], k0[fld]) < (1e-12 if precision == dg.DG_FP64 else 2e-6)
    s.close()
    s2.close()


def test_graph_replay_equals_eager_with_learned_faces_3d():
    print("case closure graph vs eager 3d mixed", flush=True)
    v0, rho0, c0 = (20.0, 10.0, -15.0), 1.225, 340.0
    m, bc = _learned_case3d(4, 3)
    bits, _ = _params(10, 12, scale=0.2, dim=3)
    out = []
    for graph in (True, False):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=v0, rho0=rho0, c0=c0, use_graph=graph,
                      steps_per_graph=8, precision=dg.DG_MIXED)
        s.set_closure(bits)
        s.add_gaussian(0.1, -0.2, 0.05, 0.05, 1.0)
        s.add_observer(0.6, 0.0, 0.0, 0.2)
        r = s.run_steps(40, 1e-6)
        out.append((s.get_state(), r.observers))
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.isfinite(out[0][0]).all() and np.abs(out[0][0][0]).max() > 0
