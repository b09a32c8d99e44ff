"""Pins of the CPU oracle against what the mathematics and the reference's own files fix.

- polynomial exactness of the weak form (volume, lift, face maps, orientation, centring)
- reflecting walls: exact for wall-compatible polynomial states (src/Mesh.cpp:911-949)
- energy decay of the Rusanov-type dissipation with sigma = +1 (src/Mesh.cpp:785-797)
- RK4 stage structure: one step equals R(dt L) with R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24
  (src/solver.cpp:488-516); forward Euler equals I + dt L (src/solver.cpp:229-230)
- residuals and observers against their closed forms (src/solver.cpp:518-573)
- the save schedule of the reference's own 2d_hilltop run (results.pvd, residuals.csv)
"""
import json
import os

import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL2, ALL3, make_case, oracle_for, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
RHO0, C0 = 1.225, 100.0


def _poly(dim, p, seed):
    """Random polynomial of total degree <= p: value and gradient callables."""
    rng = np.random.default_rng(seed)
    terms = [(a, b, c) for a in range(p + 1) for b in range(p + 1 - a) for c in range(p + 1 - a - b)
             if dim == 3 or c == 0]
    coef = rng.normal(size=len(terms))

    def val(X):
        x, y, z = X.T
        return sum(cf * x ** a * y ** b * z ** c for cf, (a, b, c) in zip(coef, terms))

    def grad(X):
        x, y, z = X.T
        g = np.zeros((len(X), 3))
        for cf, (a, b, c) in zip(coef, terms):
            if a:
                g[:, 0] += cf * a * x ** (a - 1) * y ** b * z ** c
            if b:
                g[:, 1] += cf * b * x ** a * y ** (b - 1) * z ** c
            if c:
                g[:, 2] += cf * c * x ** a * y ** b * z ** (c - 1)
        return g
    return val, grad


def _rate(grads, v0, rho0, c0):
    """-(sum_x A_x d/dx) q for the fields (p, vx, vy, vz)."""
    gp, gv = grads[0], grads[1:]
    v0 = np.asarray(v0, dtype=float)
    out = np.zeros((4, len(gp)))
    out[0] = -(gp @ v0 + rho0 * c0 * c0 * sum(gv[k][:, k] for k in range(3)))
    for k in range(3):
        out[1 + k] = -(gv[k] @ v0 + gp[:, k] / rho0)
    return out


@pytest.mark.parametrize("dim,n,p,v0", [(2, 6, 3, (-30, 30, 0)), (2, 5, 2, (10, -5, 0)), (2, 4, 5, (3, 7, 0)),
                                        (3, 2, 2, (30, 10, -5)), (3, 2, 3, (30, 0, 0))])
def test_oracle_polynomial_exactness(dim, n, p, v0):
    groups = [(ALL2 if dim == 2 else ALL3, 0)]
    m, bc, gn = make_case(dim, n, p, groups, seed=5)
    o = oracle_for(m, gn, groups, v0, RHO0, C0)
    X = m.node_coords.reshape(-1, 3)
    polys = [_poly(dim, p, 10 + f) for f in range(4)]
    u = np.stack([pv(X) for pv, _ in polys])
    grads = [pg(X) for _, pg in polys]
    if dim == 2:
        for g in grads:
            g[:, 2] = 0.0
    k = o.rhs(u, 1.0)
    expect = _rate(grads, v0, RHO0, C0)
    inner = np.repeat(np.all(m.face_side < 0, axis=1), m.Np)
    assert inner.sum() > 0
    for f in range(4):
        assert rel_l2(k[f][inner], expect[f][inner]) < 1e-10, f


@pytest.mark.parametrize("dim,n,p", [(2, 5, 3), (3, 2, 2)])
def test_oracle_reflecting_walls_exact_for_wall_compatible_states(dim, n, p):
    groups = [(ALL2 if dim == 2 else ALL3, 1)]
    m, bc, gn = make_case(dim, n, p, groups, seed=6)
    o = oracle_for(m, gn, groups, (0, 0, 0), RHO0, C0)
    X = m.node_coords.reshape(-1, 3)
    x, y, z = X.T
    pv, pg = _poly(dim, p, 3)
    if dim == 2:   # v.n = 0 on x = +-1 and y = +-1, degree <= p
        vx = (1 - x ** 2) * (0.3 + 0.7 * y)
        vy = (1 - y ** 2) * (-0.2 + 0.5 * x)
        gvx = np.stack([-2 * x * (0.3 + 0.7 * y), 0.7 * (1 - x ** 2), 0 * x], 1)
        gvy = np.stack([0.5 * (1 - y ** 2), -2 * y * (-0.2 + 0.5 * x), 0 * x], 1)
        vz, gvz = 0 * x, np.zeros((len(x), 3))
    else:
        vx, vy, vz = 0.4 * (1 - x ** 2), -0.3 * (1 - y ** 2), 0.2 * (1 - z ** 2)
        gvx = np.stack([-0.8 * x, 0 * x, 0 * x], 1)
        gvy = np.stack([0 * x, 0.6 * y, 0 * x], 1)
        gvz = np.stack([0 * x, 0 * x, -0.4 * z], 1)
    gp = pg(X)
    if dim == 2:
        gp[:, 2] = 0
    u = np.stack([pv(X), vx, vy, vz])
    k = o.rhs(u, 1.0)
    expect = _rate([gp, gvx, gvy, gvz], (0, 0, 0), RHO0, C0)
    for f in range(dim + 1):
        assert rel_l2(k[f], expect[f]) < 1e-10, f


def _energy(m, u, rho0, c0):
    ops = dg.ref_operators(m.dim, m.order)
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, m.order)
    detJ = np.abs(1.0 / np.linalg.det(pre["Jinv"][:, :m.dim, :m.dim]))
    q = np.zeros((m.K, m.Np, 4))
    q[np.arange(m.K)[:, None], pre["to_canon"], :] = u.T.reshape(m.K, m.Np, 4)
    Mq = np.einsum("ij,ejf->eif", ops["M"], q)
    quad = np.einsum("eif,eif->ef", q, Mq) * detJ[:, None]
    return 0.5 * (quad[:, 0].sum() / (rho0 * c0 * c0) + rho0 * quad[:, 1:].sum())


def test_oracle_energy_decays_with_reflecting_walls():
    groups = [(ALL2, 1)]
    m, bc, gn = make_case(2, 6, 2, groups, seed=8)
    o = oracle_for(m, gn, groups, (0, 0, 0), RHO0, C0)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(4, m.K * m.Np))
    u[1:] /= RHO0 * C0
    u[3] = 0
    dt = 2e-5
    energies = [_energy(m, u, RHO0, C0)]
    for s in range(20):
        u = o.run(u, s * dt, (s + 0.5) * dt, dt, 1e9)["u"]
        energies.append(_energy(m, u, RHO0, C0))
    e = np.array(energies)
    assert np.all(np.diff(e) <= 1e-12 * e[0])
    assert e[-1] < 0.9 * e[0]


def test_oracle_rk4_and_euler_match_amplification_polynomials():
    groups = [(0x3, 1), (0xC, 0)]
    m, bc, gn = make_case(2, 2, 1, groups, seed=9)
    v0 = (20.0, -10.0, 0.0)
    o = oracle_for(m, gn, groups, v0, RHO0, C0)
    N = 4 * m.K * m.Np
    L = np.zeros((N, N))
    for i in range(N):
        e = np.zeros(N)
        e[i] = 1.0
        L[:, i] = o.rhs(e.reshape(4, -1), 1.0).ravel()
    rng = np.random.default_rng(2)
    u0 = rng.normal(size=N)
    dt = 1e-4
    Z = dt * L
    R = np.eye(N) + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24
    u_rk = o.run(u0.reshape(4, -1), 0.0, 0.5 * dt, dt, 1e9)["u"].ravel()
    assert rel_l2(u_rk, R @ u0) < 1e-12
    u_eu = o.run(u0.reshape(4, -1), 0.0, 0.5 * dt, dt, 1e9, integrator=1)["u"].ravel()
    assert rel_l2(u_eu, u0 + Z @ u0) < 1e-12


def test_oracle_residuals_and_observers_closed_forms():
    groups = [(ALL2, 1)]
    m, bc, gn = make_case(2, 4, 2, groups, seed=4)
    o = oracle_for(m, gn, groups, (5.0, 0, 0), RHO0, C0)
    X = m.node_coords.reshape(-1, 3)
    rng = np.random.default_rng(3)
    u0 = rng.normal(size=(4, len(X)))
    obs = [(0.3, -0.2, 0.0, 0.45), (-0.5, 0.5, 0.0, 0.3)]
    r = o.run(u0, 0.0, 0.5e-5, 1e-5, 1e9, observers=obs)
    u1 = r["u"]
    exp = [np.mean((u1[0] - u0[0]) ** 2), np.mean((u0[0] / C0 ** 2 - u1[0] / C0 ** 2) ** 2),
           np.mean((u1[1] - u0[1]) ** 2), np.mean((u1[2] - u0[2]) ** 2), np.mean((u1[3] - u0[3]) ** 2)]
    np.testing.assert_allclose(r["residuals"][0], exp, rtol=1e-12)
    for i, (x, y, z, size) in enumerate(obs):
        d = np.sqrt((X[:, 0] - x) ** 2 + (X[:, 1] - y) ** 2 + (X[:, 2] - z) ** 2)
        sel = d < size
        w = 1.0 / (d[sel] ** 2 + 1e-12)
        np.testing.assert_allclose(r["observers"][0, i], (u1[:, sel] * w).sum(1) / w.sum(), rtol=1e-12)


def test_oracle_save_schedule_matches_reference_run():
    """results.pvd:3-12 lists saves at steps 0..900 and residuals.csv has 1000 rows for
    dt=1e-3, rate=0.1, tEnd=1.0 (accumulated t reaches 1.0000000000000007 > tEnd)."""
    fx = json.load(open(os.path.join(HERE, "golden", "results_pvd.json")))
    groups = [(ALL2, 1)]
    m, bc, gn = make_case(2, 1, 1, groups)
    o = oracle_for(m, gn, groups, (0, 0, 0), RHO0, 340.0)
    r = o.run(np.zeros((4, m.K * m.Np)), fx["t_start"], fx["t_end"], fx["dt"], fx["t_rate"], max_steps=2000)
    assert r["steps"] == fx["num_residual_rows"] == 1000
    assert list(r["saves"]) == [s["step"] for s in fx["saves"]]
    for s in fx["saves"]:
        assert f"{r['times'][s['step']]:.6g}" == s["timestep"]


def _flux_jacobian(n, v0, rho0, c0):
    A = np.zeros((4, 4))
    vn0 = float(np.dot(n, v0))
    A[0, 0] = vn0
    A[0, 1:] = rho0 * c0 * c0 * n
    A[1:, 0] = n / rho0
    A[1:, 1:] = vn0 * np.eye(3)
    return A


def test_oracle_absorbing_flux_is_scaled_characteristic_outflow():
    """The reference's R*K*R^-1 (src/Mesh.cpp:645-663) written through the spectral
    projectors of A.n (numpy eigendecomposition): 1/2 (vn0+c0) P_+ on the outgoing acoustic
    wave and 1/4 max(vn0, 0) P_t on the tangential (vorticity) waves."""
    import oracle_py as orc

    rng = np.random.default_rng(11)
    rho0, c0 = 1.225, 340.0
    for trial in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        s = np.cross(n, t)
        v0 = rng.normal(size=3) * 60.0
        A = _flux_jacobian(n, v0, rho0, c0)
        # symmetrised by the energy weights W = diag(1/(rho0 c0^2), rho0, rho0, rho0)
        w = np.sqrt(np.array([1.0 / (rho0 * c0 * c0), rho0, rho0, rho0]))
        lam, Q = np.linalg.eigh((w[:, None] * A) / w[None, :])
        vn0 = float(n @ v0)

        def proj(sel):
            Qs = Q[:, sel]
            return (Qs @ Qs.T) * (w[None, :] / w[:, None])

        Pp = proj(np.abs(lam - (vn0 + c0)) < 1e-6 * c0)
        Pt = proj(np.abs(lam - vn0) < 1e-6 * c0)
        M = 0.5 * (vn0 + c0) * Pp + 0.25 * max(vn0, 0.0) * Pt
        for _ in range(3):
            q = rng.normal(size=4)
            out = orc.bc_flux(0, n, t, s, q, v0=v0, rho0=rho0, c0=c0)
            np.testing.assert_allclose(out, M @ q, rtol=1e-10, atol=1e-10 * np.abs(M @ q).max())


def test_oracle_reflecting_flux_is_wall_flux():
    import oracle_py as orc

    rng = np.random.default_rng(12)
    for _ in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v0 = rng.normal(size=3) * 30.0
        q = rng.normal(size=4)
        qw = q.copy()
        qw[1:] -= np.dot(n, q[1:]) * n
        out = orc.bc_flux(1, n, n, n, q, v0=v0, rho0=1.225, c0=340.0)
        np.testing.assert_allclose(out, _flux_jacobian(n, v0, 1.225, 340.0) @ qw, rtol=1e-12, atol=1e-9)
