"""CPU checks of the learned-closure oracle (pinned by finite differences, torch's BF16
conversion and invariances) and of the library's host-side closure logic."""
import numpy as np
import pytest

import closure_oracle as co
import dgfem_amd as dg


def _rand_params(nfp, seed, scale=0.5):
    nin, nout = co.sizes(nfp)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, dg.closure_num_params(nfp)).astype(np.float32)
    return p, nin, nout


def test_oracle_gradient_matches_finite_differences():
    nfp = 3
    p32, nin, nout = _rand_params(nfp, 1)
    p = p32.astype(np.float64)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (7, nin))
    Y = rng.normal(0, 1, (7, nout))
    L, g = co.loss_grad(p, X, Y)
    h = 1e-6
    idx = np.concatenate([rng.choice(p.size, 60, replace=False), [0, nin * 64, p.size - 1, p.size - nout - 1]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (co.loss_grad(p + e, X, Y)[0] - co.loss_grad(p - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_matches_torch_and_library():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000)]).astype(
        np.float32)
    # exact ties between two BF16 neighbours, both parities
    base = rng.integers(0x3F00, 0x4100, 500).astype(np.uint32) << 16
    ties = (base | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().astype(np.uint16)
    assert np.array_equal(co.f32_to_bf16(x), ref)
    assert np.array_equal(dg.to_bf16(x), ref)
    assert np.array_equal(co.bf16_to_f64(ref), torch.from_numpy(ref.view(np.int16)).view(torch.bfloat16).double().numpy())


def _random_faces(m, nfp, seed, nf=4):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, m)
    normals = np.stack([np.cos(ang), np.sin(ang), np.zeros(m)], axis=1)
    qi = rng.normal(0, 1, (m, nfp, nf))
    qi[..., 3] = 0.0
    qe = qi + rng.normal(0, 0.1, (m, nfp, nf))
    qe[..., 3] = 0.0
    return qi, qe, normals


def test_features_invariances_and_hand_case():
    rho0, c0, v0 = 1.2, 340.0, np.array([30.0, -20.0, 0.0])
    qi, qe, nrm = _random_faces(50, 4, 5)
    X, s = co.features(qi, nrm, v0, rho0, c0)
    # rotation of the whole configuration by theta leaves the features unchanged
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    qr = qi.copy()
    qr[..., 1:3] = qi[..., 1:3] @ R.T
    nr = nrm.copy()
    nr[:, :2] = nrm[:, :2] @ R.T
    Xr, _ = co.features(qr, nr, np.r_[R @ v0[:2], 0.0], rho0, c0)
    np.testing.assert_allclose(Xr, X, atol=1e-12)
    # scaling the trace scales s and leaves x unchanged
    Xs, ss = co.features(3.5 * qi, nrm, v0, rho0, c0)
    np.testing.assert_allclose(Xs, X, atol=1e-12)
    np.testing.assert_allclose(ss, 3.5 * s, rtol=1e-12)
    # hand case: n = (1, 0), p = rho0 c0, v = (2, 0) at both nodes: P = 1, N = 2, T = 0,
    # w_out = 3, s = 3
    q = np.zeros((1, 2, 4))
    q[..., 0] = rho0 * c0
    q[..., 1] = 2.0
    Xh, sh = co.features(q, np.array([[1.0, 0.0, 0.0]]), v0, rho0, c0)
    assert np.isclose(sh[0], 3.0)
    np.testing.assert_allclose(Xh[0], [1, 1, 0, 0, 30 / c0, -20 / c0])
    # target: incoming characteristic of the neighbour trace, P - N = 1 - 2
    np.testing.assert_allclose(co.targets(q, np.array([[1.0, 0.0, 0.0]]), rho0, c0, sh), [[-1 / 3, -1 / 3]])


def _lf_flux(q_m, q_p, n, v0, rho0, c0):
    """Lax-Friedrichs face flux in physical variables (the solver's interior flux)."""
    def F(q):
        vn0 = n[:, 0:1] * v0[0] + n[:, 1:2] * v0[1]
        vn = q[..., 1] * n[:, 0:1] + q[..., 2] * n[:, 1:2]
        out = np.zeros_like(q)
        out[..., 0] = vn0 * q[..., 0] + rho0 * c0 * c0 * vn
        for r in range(2):
            out[..., 1 + r] = vn0 * q[..., 1 + r] + n[:, r:r + 1] * q[..., 0] / rho0
        return out
    return 0.5 * (F(q_m) + F(q_p)) + 0.5 * c0 * (q_m - q_p)


@pytest.mark.parametrize("vmag", [0.0, 60.0, 120.0])
def test_exterior_state_gives_the_upwind_flux(vmag):
    """q+ built from a prescribed incoming characteristic must turn the Lax-Friedrichs flux
    into the upwind flux: outgoing (vn0 + c0) w_out-, incoming (vn0 - c0) w_hat, vorticity
    vn0 T- on outflow and 0 on inflow."""
    rho0, c0 = 1.2, 340.0
    rng = np.random.default_rng(int(vmag) + 1)
    qi, _, nrm = _random_faces(40, 5, 9)
    ang = rng.uniform(0, 2 * np.pi)
    v0 = np.array([vmag * np.cos(ang), vmag * np.sin(ang), 0.0])
    w_hat = rng.normal(0, 1, (40, 5))
    qp = co.exterior_from_incoming(qi, nrm, w_hat, v0, rho0, c0)
    got = _lf_flux(qi, qp, nrm[:, :2], v0, rho0, c0)
    n, t = co.frame(nrm)
    t = t[:, 0]
    wo, _, T = co.characteristics(qi, nrm, rho0, c0)
    T = T[:, 0]
    vn0 = (n @ v0[:2])[:, None]
    FP = 0.5 * ((vn0 + c0) * wo + (vn0 - c0) * w_hat)
    FN = 0.5 * ((vn0 + c0) * wo - (vn0 - c0) * w_hat)
    FT = np.where(vn0 >= 0, vn0 * T, 0.0)
    want = np.zeros_like(got)
    want[..., 0] = rho0 * c0 * FP
    want[..., 1] = FN * n[:, 0:1] + FT * t[:, 0:1]
    want[..., 2] = FN * n[:, 1:2] + FT * t[:, 1:2]
    np.testing.assert_allclose(got[..., :3], want[..., :3], atol=1e-9 * np.abs(want).max())
    # the zero network is this exterior state with w_hat = 0
    zero = np.zeros(dg.closure_num_params(5))
    np.testing.assert_allclose(co.exterior(qi, nrm, zero, v0, rho0, c0),
                               co.exterior_from_incoming(qi, nrm, 0 * w_hat, v0, rho0, c0), atol=0)


def test_library_features_match_oracle():
    rho0, c0, v0 = 1.225, 340.0, (50.0, 10.0, 0.0)
    for nfp in (2, 5):
        qi, qe, nrm = _random_faces(200, nfp, 11 + nfp)
        traces = np.stack([qi, qe], axis=1)
        X, Y, S = dg.closure_features(traces, nrm, v0, rho0, c0)
        Xo, s = co.features(qi, nrm, v0, rho0, c0)
        Yo = co.targets(qe, nrm, rho0, c0, s)
        np.testing.assert_allclose(S, s, rtol=2e-6)
        np.testing.assert_allclose(X, Xo, rtol=2e-6, atol=2e-6)
        np.testing.assert_allclose(Y, Yo, rtol=2e-5, atol=2e-6)


def test_learned_faces_numbering_in_precompute():
    m = dg.box_mesh(2, 4, 2, jitter=0.1, seed=3)
    bc = np.full((m.K, 3), dg.DG_BC_LEARNED, dtype=np.int8)
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, 2, bc)
    nbr = pre["nbr"]
    codes = nbr[nbr < 0]
    assert codes.size == 16  # 4 sides x 4 edges
    assert sorted(codes.tolist()) == [-3 - i for i in range(16)][::-1]
    flat = nbr.ravel()
    order = [-3 - v for v in flat if v <= -3]
    assert order == list(range(16))  # numbered in (element, local face) order


def test_learned_closure_rejected_on_high_order_tets():
    m = dg.box_mesh(3, 1, 4, seed=1)
    bc = np.full((m.K, 4), dg.DG_BC_LEARNED, dtype=np.int8)
    with pytest.raises(dg.DGError) as e:
        dg.Solver(m.node_coords, m.element_vertices, 4, bc)
    assert e.value.code == dg.DG_ERR_UNSUPPORTED and "order <= 3" in str(e.value)


def test_learned_faces_numbered_on_tets():
    m = dg.box_mesh(3, 2, 2, seed=3)
    bc = np.where(m.face_side >= 0, dg.DG_BC_LEARNED, 0).astype(np.int8)
    nbr = dg.mesh_precompute(m.node_coords, m.element_vertices, 2, bc)["nbr"]
    flat = nbr.ravel()
    assert [-3 - v for v in flat if v <= -3] == list(range(6 * 2 * 2 * 2))  # 6 sides x 4 squares x 2 triangles
    assert (flat >= -2).sum() + (flat <= -3).sum() == flat.size and not np.any((flat == -1) | (flat == -2))


# ---- 3D faces ----

def _random_faces3d(m, nfp, seed, nf=4):
    rng = np.random.default_rng(seed)
    n = rng.normal(0, 1, (m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    qi = rng.normal(0, 1, (m, nfp, nf))
    qe = qi + rng.normal(0, 0.1, (m, nfp, nf))
    return qi, qe, n


def _lf_flux3d(q_m, q_p, n, v0, rho0, c0):
    """Lax-Friedrichs face flux in physical variables, 3D (the solver's interior flux)."""
    def F(q):
        vn0 = (n @ v0)[:, None]
        vn = np.einsum("mjc,mc->mj", q[..., 1:4], n)
        out = np.zeros_like(q)
        out[..., 0] = vn0 * q[..., 0] + rho0 * c0 * c0 * vn
        for r in range(3):
            out[..., 1 + r] = vn0 * q[..., 1 + r] + n[:, r:r + 1] * q[..., 0] / rho0
        return out
    return 0.5 * (F(q_m) + F(q_p)) + 0.5 * c0 * (q_m - q_p)


@pytest.mark.parametrize("vmag", [0.0, 60.0, 120.0])
def test_exterior_state_gives_the_upwind_flux_3d(vmag):
    """3D: the exterior state turns the Lax-Friedrichs flux into the upwind flux with the
    prescribed incoming characteristic, outgoing (vn0 + c0) w_out-, and the tangential velocity
    (both components) carried out on outflow and zero on inflow. The expected flux is built
    from the eigenvectors of A.n written independently of the oracle's frame."""
    rho0, c0 = 1.2, 340.0
    rng = np.random.default_rng(int(vmag) + 7)
    qi, _, nrm = _random_faces3d(40, 10, 19)
    d = rng.normal(0, 1, 3)
    v0 = vmag * d / np.linalg.norm(d)
    w_hat = rng.normal(0, 1, (40, 10))
    qp = co.exterior_from_incoming(qi, nrm, w_hat, v0, rho0, c0, dim=3)
    got = _lf_flux3d(qi, qp, nrm, v0, rho0, c0)
    vn0 = (nrm @ v0)[:, None]
    v = qi[..., 1:4]
    vn = np.einsum("mjc,mc->mj", v, nrm)
    wo = qi[..., 0] / (rho0 * c0) + vn
    vt = v - vn[..., None] * nrm[:, None, :]                 # tangential velocity, frame-free
    FP = 0.5 * ((vn0 + c0) * wo + (vn0 - c0) * w_hat)
    FN = 0.5 * ((vn0 + c0) * wo - (vn0 - c0) * w_hat)
    FT = np.where((vn0 >= 0)[..., None], vn0[..., None] * vt, 0.0)
    want = np.zeros_like(got)
    want[..., 0] = rho0 * c0 * FP
    want[..., 1:4] = FN[..., None] * nrm[:, None, :] + FT
    np.testing.assert_allclose(got, want, atol=1e-9 * np.abs(want).max())


def test_features_rotation_invariance_3d():
    """Rotating traces, normals and mean flow together leaves the 3D features unchanged (the
    tangent frame follows the tangential mean flow); the scale is homogeneous of degree one."""
    rho0, c0 = 1.225, 340.0
    qi, _, nrm = _random_faces3d(30, 10, 23)
    v0 = np.array([40.0, -25.0, 60.0])
    X, s = co.features(qi, nrm, v0, rho0, c0, dim=3)
    A = np.random.default_rng(5).normal(0, 1, (3, 3))
    R, _ = np.linalg.qr(A)
    qr = qi.copy()
    qr[..., 1:4] = qi[..., 1:4] @ R.T
    Xr, sr = co.features(qr, nrm @ R.T, R @ v0, rho0, c0, dim=3)
    np.testing.assert_allclose(Xr, X, atol=1e-12)
    np.testing.assert_allclose(sr, s, rtol=1e-13)
    X2, s2 = co.features(2.5 * qi, nrm, v0, rho0, c0, dim=3)
    np.testing.assert_allclose(X2, X, atol=1e-12)
    np.testing.assert_allclose(s2, 2.5 * s, rtol=1e-13)
    # no tangential mean flow: the frame follows the coordinate axes; the features still see
    # the same tangential speed |v_t| per node
    Xa, sa = co.features(qi, nrm, (nrm[0] * 50.0), rho0, c0, dim=3)
    T2 = Xa[:, 10:20] ** 2 + Xa[:, 20:30] ** 2
    vt = qi[..., 1:4] - np.einsum("mjc,mc->mj", qi[..., 1:4], nrm)[..., None] * nrm[:, None, :]
    np.testing.assert_allclose(T2 * sa[:, None] ** 2, (vt ** 2).sum(-1), rtol=1e-10)


def test_library_features_match_oracle_3d():
    rho0, c0 = 1.225, 340.0
    for v0, nfp in (((50.0, 10.0, -30.0), 10), ((0.0, 0.0, 0.0), 6), ((20.0, 0.0, 0.0), 3)):
        qi, qe, nrm = _random_faces3d(200, nfp, 31 + nfp)
        nrm[:5] = np.eye(3)[[0, 1, 2, 0, 1]] * np.array([1, -1, 1, -1, 1])[:, None]  # axis-aligned faces
        traces = np.stack([qi, qe], axis=1)
        X, Y, S = dg.closure_features(traces, nrm, v0, rho0, c0, dim=3)
        Xo, s = co.features(qi, nrm, v0, rho0, c0, dim=3)
        Yo = co.targets(qe, nrm, rho0, c0, s, v0, dim=3)
        assert X.shape == (200, 3 * nfp + 2)
        np.testing.assert_allclose(S, s, rtol=2e-6)
        np.testing.assert_allclose(X, Xo, rtol=2e-6, atol=2e-6)
        np.testing.assert_allclose(Y, Yo, rtol=2e-5, atol=2e-6)
    assert dg.closure_num_params(10, dim=3) == 64 * 32 + 64 + 64 * 64 + 64 + 10 * 64 + 10


def test_saved_network_round_trip(tmp_path):
    """The split pipeline's network file: FP32 master weights, the output blend and the case,
    read back without unpickling."""
    import closure as cl
    case = cl.ClosureCase(dim=3, n_trunc=8, factor=2, order=3, jitter=0.15, steps=400)
    master = dg.closure_init(cl.face_nodes(case), seed=3, out_scale=0.5, dim=3)
    path = str(tmp_path / "net.npz")
    cl.save_network(path, master, case, 0.25, train_runs=3)
    m2, alpha, meta = cl.load_network(path)
    assert np.array_equal(m2, master) and m2.dtype == np.float32 and alpha == 0.25
    assert meta["case"]["n_trunc"] == 8 and meta["case"]["steps"] == 400 and meta["train_runs"] == 3
    cl.save_network(path, master, case, None)
    assert cl.load_network(path)[1] is None
