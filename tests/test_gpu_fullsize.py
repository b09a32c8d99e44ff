"""Parity at the configured sizes.

2D LEE case (the CPU-runnable configuration): [-1,1]^2, 96x96 quads split along seeded
diagonals, interior vertices jittered by 0.2h (18,432 triangles), p=3, v0=(-30,30), rho0=1.225,
c0=100, Gaussian pulse exp(-r^2/0.1) at the origin, reflecting walls, 500 RK4 steps, observers at
(+-0.8, 0). With the configured dt = 1e-5 this mesh is beyond the RK4 stability limit: the
oracle's solution grows from |p| ~ 1 to ~1e60 in 500 steps (checked here), so it is run at its
CFL-limited dt = r_min / ((c0 + |v0|)(2p + 1)) instead. The GPU FP64 path must match the oracle
to 1e-10 relative L2 in every field and observer series; the mixed path must stay within 1e-6
of FP64 pressure.

3D APE case at full size (32^3x6 jittered Kuhn tets, K=196,608, p=3, 15.7M DOF, absorbing):
one right-hand side against the oracle (1e-12), two RK4 steps against the oracle (1e-11,
with the residual and observer outputs), and linearity of the FP64 right-hand side.
"""
import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL2, ALL3, gaussian_state, make_case, oracle_for, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

RHO0 = 1.225


@pytest.fixture(scope="module")
def c1():
    groups = [(ALL2, dg.DG_BC_REFLECTING)]
    m, bc, gn = make_case(2, 96, 3, groups, jitter=0.2, seed=1)
    assert m.K == 18432
    return m, bc, gn, groups


def _c1_dt(m, bc):
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(-30.0, 30.0, 0.0), rho0=RHO0, c0=100.0)
    r_min = s.info().min_inradius
    s.close()
    return r_min / ((100.0 + np.hypot(30.0, 30.0)) * 7)


def test_c1_configured_dt_is_unstable(c1):
    """The configured dt = 1e-5 on this mesh: the oracle's pressure grows by > 1e10 within
    200 steps (the reference algorithm diverges), which is why the parity runs use CFL dt."""
    print("case c1 configured dt stability check (oracle only)", flush=True)
    m, bc, gn, groups = c1
    o = oracle_for(m, gn, groups, (-30.0, 30.0, 0.0), RHO0, 100.0)
    u0 = gaussian_state(m, size=0.1)
    r = o.run(u0, 0.0, 199.5e-5, 1e-5, 1e9)
    assert np.abs(r["u"][0]).max() > 1e10 * np.abs(u0[0]).max()
    assert 1e-5 > 2 * _c1_dt(m, bc)


def _c1_solver(m, bc, precision):
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(-30.0, 30.0, 0.0), rho0=RHO0, c0=100.0,
                  precision=precision)
    s.set_state(gaussian_state(m, size=0.1))
    for x in (0.8, -0.8):
        assert s.add_observer(x, 0.0, 0.0, 0.05) > 0
    return s


def test_c1_full_configuration_matches_oracle(c1):
    print("case c1 full 96x96x2 p3 500 steps fp64 vs oracle", flush=True)
    m, bc, gn, groups = c1
    o = oracle_for(m, gn, groups, (-30.0, 30.0, 0.0), RHO0, 100.0)
    u0 = gaussian_state(m, size=0.1)
    obs = [(0.8, 0.0, 0.0, 0.05), (-0.8, 0.0, 0.0, 0.05)]
    dt = _c1_dt(m, bc)
    ref = o.run(u0, 0.0, 499.5 * dt, dt, 1e9, observers=obs)
    assert ref["steps"] == 500
    assert np.abs(ref["u"][0]).max() < 2.0  # stable at this dt
    s = _c1_solver(m, bc, dg.DG_FP64)
    r = s.run_steps(500, dt)
    u = s.get_state()
    for f in range(3):
        e = rel_l2(u[f], ref["u"][f])
        print(f"field {f}: rel L2 {e:.2e}")
        assert e < 1e-10
    for j in range(2):
        for k in range(3):
            assert rel_l2(r.observers[:, j, k], ref["observers"][:, j, k]) < 1e-10
    np.testing.assert_allclose(r.residuals[:, :4], ref["residuals"][:, :4], rtol=1e-8, atol=0)


def test_c1_mixed_precision_within_1e6_of_fp64(c1):
    print("case c1 full 500 steps mixed vs fp64", flush=True)
    m, bc, _, _ = c1
    out = {}
    dt = _c1_dt(m, bc)
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        s = _c1_solver(m, bc, prec)
        r = s.run_steps(500, dt)
        out[prec] = (s.get_state(), r.observers)
    e = rel_l2(out[dg.DG_MIXED][0][0], out[dg.DG_FP64][0][0])
    eo = rel_l2(out[dg.DG_MIXED][1][:, :, 0], out[dg.DG_FP64][1][:, :, 0])
    print(f"mixed rel L2 p {e:.2e} observers {eo:.2e}")
    assert e < 1e-6 and eo < 1e-6


@pytest.fixture(scope="module")
def c2():
    groups = [(ALL3, dg.DG_BC_ABSORBING)]
    m, bc, gn = make_case(3, 32, 3, groups, jitter=0.15, seed=2)
    assert m.K == 196608
    return m, bc, gn, groups


def test_c2_full_size_rhs_matches_oracle(c2):
    print("case c2 full 32^3x6 p3 rhs fp64 vs oracle", flush=True)
    m, bc, gn, groups = c2
    o = oracle_for(m, gn, groups, (30.0, 0.0, 0.0), RHO0, 340.0)
    u = smooth_state(m, seed=3)
    k_ref = o.rhs(u, 2.5e-6)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=340.0)
    k = s.rhs(u, 2.5e-6)
    for f in range(4):
        e = rel_l2(k[f], k_ref[f])
        print(f"field {f}: rel L2 {e:.2e}")
        assert e < 1e-12
    # linearity of the right-hand side (every flux and boundary branch is linear in q)
    u2 = gaussian_state(m, 0.2, -0.1, 0.3, size=0.05)
    a, b = 0.7, -1.9
    kl = s.rhs(a * u + b * u2, 2.5e-6)
    k2 = s.rhs(u2, 2.5e-6)
    for f in range(4):
        assert rel_l2(kl[f], a * k[f] + b * k2[f]) < 1e-12


def test_c2_full_size_two_rk4_steps_match_oracle(c2):
    print("case c2 full 32^3x6 p3 two rk4 steps with monopole vs oracle", flush=True)
    m, bc, gn, groups = c2
    o = oracle_for(m, gn, groups, (30.0, 0.0, 0.0), RHO0, 340.0)
    u0 = smooth_state(m, seed=4, noise=0.0)
    src = [dict(kind=0, x=0.0, y=0.0, z=0.0, size=0.1, amp=0.1, freq=200.0, phase=0.0, duration=1e9)]
    obs = [(0.5, 0.0, 0.0, 0.1), (0.0, -0.5, 0.25, 0.1)]
    dt = 2.5e-6
    ref = o.run(u0, 0.0, 1.5 * dt, dt, 1e9, sources=src, observers=obs)
    assert ref["steps"] == 2
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=340.0)
    s.set_state(u0)
    s.add_monopole(0.0, 0.0, 0.0, 0.1, 0.1, 200.0, 0.0, 1e9)
    for ob in obs:
        s.add_observer(*ob)
    r = s.run_steps(2, dt)
    u = s.get_state()
    for f in range(4):
        assert rel_l2(u[f], ref["u"][f]) < 1e-11
    np.testing.assert_allclose(r.residuals[:, :4], ref["residuals"][:, :4], rtol=1e-9, atol=1e-30)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-10, atol=1e-14)


def _c2_solver(m, bc, precision):
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=340.0,
                  precision=precision, steps_per_graph=20)
    assert s.add_monopole(0.0, 0.0, 0.0, 0.1, 0.1, 200.0, 0.0, 1e9) > 0
    assert s.add_observer(0.5, 0.0, 0.0, 0.1) > 0
    assert s.add_observer(0.0, -0.5, 0.25, 0.1) > 0
    return s


@pytest.mark.timeout(600)
def test_c2_2000_steps_mixed_within_1e6_of_fp64(c2):
    """The headline condition at the configured size and length: the 3D APE case (196,608 tets,
    p=3, monopole at the origin, absorbing, CFL-limited dt) from a zero state, 2000 RK4 steps in
    FP64 and in mixed precision. Pressure field and observer pressure series: rel L2 <= 1e-6."""
    print("case c2 full 32^3x6 p3 2000 steps mixed vs fp64", flush=True)
    m, bc, _, _ = c2
    probe = _c2_solver(m, bc, dg.DG_FP64)
    dt = probe.info().min_inradius / ((340.0 + 30.0) * 7)
    probe.close()
    out = {}
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        s = _c2_solver(m, bc, prec)
        r = s.run_steps(2000, dt)
        assert r.steps == 2000
        out[prec] = (s.get_state(), r.observers)
        s.close()
        print(f"precision {prec}: {r.device_ms / 2000:.4f} ms/step", flush=True)
    u64, o64 = out[dg.DG_FP64]
    umx, omx = out[dg.DG_MIXED]
    e = [rel_l2(umx[f], u64[f]) for f in range(4)]
    eo = rel_l2(omx[:, :, 0], o64[:, :, 0])
    print(f"rel L2 mixed vs fp64 after 2000 steps: fields {['%.2e' % x for x in e]}, observers p {eo:.2e}")
    # the wave has filled the box: the pressure is nonzero at the far corner region
    X = m.node_coords.reshape(-1, 3)
    far = np.linalg.norm(X, axis=1) > 0.9
    assert np.abs(u64[0][far]).max() > 1e-3 * np.abs(u64[0]).max()
    assert e[0] <= 1e-6 and eo <= 1e-6
    assert max(e[1:]) <= 1e-5
