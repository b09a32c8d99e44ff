"""GPU parity of the HIP path against the CPU oracle (the reference algorithm restated).

FP64 bar: 1e-10 relative L2 per field (summation-order differences only); single RHS
evaluations 1e-12. Mixed precision bar: 1e-6 relative L2 in pressure against the GPU FP64
path.
"""
import os

import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL2, ALL3, gaussian_state, make_case, oracle_for, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

RHO0 = 1.225


def matrix_core_built(dim, p):
    """Shapes with FP32/FP64 matrix-core stage kernels: triangles, tetrahedra of order <= 3."""
    return dim == 2 or p <= 3


def kernel_built(kernel, dim, p):
    """valu everywhere; vm (volume on the matrix cores) on tetrahedra of order 2 and 3; mm and
    wave on triangles and tetrahedra of order <= 3."""
    if kernel == "valu":
        return True
    if kernel == "vm":
        return dim == 3 and p in (2, 3)
    return matrix_core_built(dim, p)


@pytest.mark.parametrize("dim,n,p,groups,v0,c0", [
    (2, 4, 1, [(ALL2, 1)], (-30, 30, 0), 100.0),
    (2, 4, 2, [(0x3, 1), (0xC, 0)], (-30, 30, 0), 100.0),
    (2, 4, 3, [(ALL2, 1)], (-30, 30, 0), 100.0),
    (2, 3, 4, [(ALL2, 0)], (10, -20, 0), 100.0),
    (2, 3, 5, [(ALL2, 0)], (10, -20, 0), 100.0),
    (3, 2, 1, [(ALL3, 0)], (30, 0, 0), 340.0),
    (3, 2, 2, [(0x10, 1), (0x2F, 0)], (30, 10, -5), 340.0),
    (3, 2, 3, [(ALL3, 0)], (30, 0, 0), 340.0),
    (3, 2, 4, [(0x10, 1), (0x2F, 0)], (30, 0, 0), 340.0),
    (3, 1, 5, [(ALL3, 0)], (30, 0, 0), 340.0),
])
@pytest.mark.parametrize("kernel", ["valu", "vm", "mm", "wave"])
def test_rhs_matches_oracle(dim, n, p, groups, v0, c0, kernel, monkeypatch):
    """The matrix-core variants are built for some shapes only (kernel_built): on the others an
    explicit request must fail dg_create with DG_ERR_UNSUPPORTED."""
    print(f"case rhs dim={dim} n={n} p={p} kernel={kernel}", flush=True)
    monkeypatch.setenv("DGFEM_KERNEL", kernel)
    m, bc, gn = make_case(dim, n, p, groups, seed=7)
    if not kernel_built(kernel, dim, p):
        with pytest.raises(dg.DGError) as ei:
            dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=c0)
        assert ei.value.code == dg.DG_ERR_UNSUPPORTED and f"DGFEM_KERNEL={kernel}" in str(ei.value)
        return
    o = oracle_for(m, gn, groups, v0, RHO0, c0)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=c0)
    assert dg.KERNEL_NAMES[s.info().stage_kernel] == kernel
    u = smooth_state(m)
    k_ref = o.rhs(u, 1e-4)
    k_gpu = s.rhs(u, 1e-4)
    for f in range(4):
        if np.linalg.norm(k_ref[f]) > 0:
            assert rel_l2(k_gpu[f], k_ref[f]) < 1e-12, (f, rel_l2(k_gpu[f], k_ref[f]))
        else:
            assert np.abs(k_gpu[f]).max() < 1e-14


def test_rk4_2d_reflecting_matches_oracle():
    """Reduced 2D LEE case (24x24 jittered triangles, p=3, v0=(-30,30), reflecting walls,
    Gaussian pulse, dt=1e-5, 500 steps, two observers)."""
    print("case rk4 2d reflecting 24x24 p3 500 steps", flush=True)
    groups = [(ALL2, 1)]
    v0 = (-30.0, 30.0, 0.0)
    m, bc, gn = make_case(2, 24, 3, groups, seed=1)
    o = oracle_for(m, gn, groups, v0, RHO0, 100.0)
    u0 = gaussian_state(m, size=0.1)
    obs = [(0.8, 0.0, 0.0, 0.05), (-0.8, 0.0, 0.0, 0.05)]
    dt, steps = 1e-5, 500
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 1e9, observers=obs)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=v0, rho0=RHO0, c0=100.0)
    s.set_state(u0)
    for ob in obs:
        assert s.add_observer(*ob) > 0
    r = s.run(0.0, (steps - 0.5) * dt, dt, 1e9)
    assert r.steps == ref["steps"] == steps
    u = s.get_state()
    for f in range(3):
        assert rel_l2(u[f], ref["u"][f]) < 1e-10, (f, rel_l2(u[f], ref["u"][f]))
    np.testing.assert_allclose(r.times, ref["times"], rtol=0, atol=0)
    np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("kernel", ["valu", "vm", "mm", "wave"])
def test_rk4_3d_absorbing_monopole_matches_oracle(kernel, monkeypatch):
    """Reduced 3D case: 4^3 x 6 jittered Kuhn tets, p=3, v0=(30,0,0), c0=340, all
    absorbing, monopole at the origin."""
    print(f"case rk4 3d absorbing monopole n4 p3 kernel={kernel}", flush=True)
    monkeypatch.setenv("DGFEM_KERNEL", kernel)
    groups = [(ALL3, 0)]
    v0 = (30.0, 0.0, 0.0)
    m, bc, gn = make_case(3, 4, 3, groups, jitter=0.15, seed=2)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    src = dict(kind=0, x=0.0, y=0.0, z=0.0, size=0.3, amp=0.1, freq=200.0, phase=0.0, duration=1.0)
    obs = [(0.5, 0.1, -0.2, 0.3)]
    dt, steps = 2e-5, 100
    u0 = np.zeros((4, m.K * m.Np))
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 1e9, sources=[src], observers=obs)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=v0, rho0=RHO0, c0=340.0)
    assert dg.KERNEL_NAMES[s.info().stage_kernel] == kernel
    s.set_state(u0)
    assert s.add_monopole(0, 0, 0, 0.3, 0.1, 200.0, 0.0, 1.0) > 0
    s.add_observer(*obs[0])
    r = s.run(0.0, (steps - 0.5) * dt, dt, 1e9)
    u = s.get_state()
    for f in range(4):
        assert rel_l2(u[f], ref["u"][f]) < 1e-10, (f, rel_l2(u[f], ref["u"][f]))
    np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-8, atol=1e-14)


def test_overlapping_sources_match_oracle():
    """Sources applied in definition order (src/solver.cpp:455-483): where sources overlap, the
    last active one sets the node. Three overlapping monopoles whose durations end at different
    steps and a table source that is active on every third step only, 60 RK4 steps in batches of
    7 (sources of a batch's first step and of later steps are written by different kernels)."""
    print("case overlapping sources 3d n4 p2", flush=True)
    groups = [(ALL3, 0)]
    v0 = (30.0, 0.0, 0.0)
    m, bc, gn = make_case(3, 4, 2, groups, jitter=0.15, seed=6)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    dt, steps = 2e-5, 60
    mono = [(0.0, 0.0, 0.0, 0.45, 0.1, 200.0, 0.0, 1.0), (0.1, 0.0, 0.0, 0.35, 0.2, 350.0, 0.5, 30.5 * dt),
            (-0.05, 0.05, 0.0, 0.25, 0.05, 500.0, 1.0, 10.5 * dt)]

    def table(t):
        k = int(round(t / dt))
        return 0.03 * np.cos(900.0 * t) if k % 3 == 0 else None

    src = [dict(kind=0, x=x, y=y, z=z, size=r, amp=a, freq=f, phase=ph, duration=d) for x, y, z, r, a, f, ph, d in mono]
    src.append(dict(kind=1, x=0.05, y=0.0, z=0.05, size=0.3))
    u0 = gaussian_state(m, size=0.05)
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 25 * dt, sources=src, table_fn=lambda i, t: table(t),
                observers=[(0.0, 0.0, 0.0, 0.3)])
    s = dg.Solver(m.node_coords, m.element_vertices, 2, bc, v0=v0, rho0=RHO0, c0=340.0, steps_per_graph=7)
    s.set_state(u0)
    counts = [s.add_monopole(*mp) for mp in mono] + [s.add_table_source(0.05, 0.0, 0.05, 0.3, table)]
    assert min(counts) > 0
    s.add_observer(0.0, 0.0, 0.0, 0.3)
    r = s.run(0.0, (steps - 0.5) * dt, dt, 25 * dt)
    u = s.get_state()
    for f in range(4):
        assert rel_l2(u[f], ref["u"][f]) < 1e-10, (f, rel_l2(u[f], ref["u"][f]))
    np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-8, atol=1e-14)


def test_more_observers_than_post_threads():
    """1,100 observers (more than the 1,024 threads of the end-of-step kernel), some of them
    sharing nodes, 12 RK4 steps against the oracle."""
    print("case 1100 observers 3d n3 p2", flush=True)
    groups = [(ALL3, 0)]
    v0 = (30.0, 0.0, 0.0)
    m, bc, gn = make_case(3, 3, 2, groups, jitter=0.15, seed=8)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    rng = np.random.default_rng(3)
    obs = [(float(x), float(y), float(z), float(r)) for (x, y, z), r in
           zip(rng.uniform(-0.8, 0.8, (1100, 3)), rng.uniform(0.25, 0.5, 1100))]
    dt, steps = 2e-5, 12
    u0 = gaussian_state(m, size=0.05)
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 1e9, observers=obs)
    s = dg.Solver(m.node_coords, m.element_vertices, 2, bc, v0=v0, rho0=RHO0, c0=340.0, steps_per_graph=5)
    s.set_state(u0)
    counts = [s.add_observer(*ob) for ob in obs]
    assert min(counts) > 0
    r = s.run(0.0, (steps - 0.5) * dt, dt, 1e9)
    assert r.observers.shape == (steps, 1100, 4)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-8, atol=1e-14)


def test_euler_matches_oracle():
    print("case euler 2d", flush=True)
    groups = [(0x3, 1), (0xC, 0)]
    v0 = (20.0, -10.0, 0.0)
    m, bc, gn = make_case(2, 8, 2, groups, seed=4)
    o = oracle_for(m, gn, groups, v0, RHO0, 100.0)
    u0 = gaussian_state(m, 0.1, -0.2, 0, 0.05)
    dt, steps = 2e-6, 40
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 1e9, integrator=1)
    s = dg.Solver(m.node_coords, m.element_vertices, 2, bc, v0=v0, rho0=RHO0, c0=100.0)
    s.set_state(u0)
    r = s.run(0.0, (steps - 0.5) * dt, dt, 1e9, integrator=dg.DG_EULER)
    u = s.get_state()
    for f in range(3):
        assert rel_l2(u[f], ref["u"][f]) < 1e-10
    np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)


@pytest.mark.parametrize("kernel", ["valu", "vm", "mm"])
def test_euler_3d_with_sources_matches_oracle(kernel, monkeypatch):
    """Forward Euler on tetrahedra (p = 3, reflecting floor, absorbing elsewhere) with a monopole
    and an observer, 30 steps in batches of 7, against the oracle; the mixed-precision Euler path
    within 1e-6 of FP64."""
    print(f"case euler 3d p3 kernel={kernel}", flush=True)
    monkeypatch.setenv("DGFEM_KERNEL", kernel)
    groups = [(0x10, 1), (0x2F, 0)]
    v0 = (25.0, -10.0, 5.0)
    m, bc, gn = make_case(3, 4, 3, groups, jitter=0.15, seed=15)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    u0 = gaussian_state(m, 0.1, -0.2, 0.1, 0.03)
    src = dict(kind=0, x=-0.2, y=0.1, z=0.0, size=0.3, amp=0.05, freq=400.0, phase=0.3, duration=1.0)
    obs = [(0.3, 0.3, -0.2, 0.3)]
    dt, steps = 1e-6, 30
    ref = o.run(u0, 0.0, (steps - 0.5) * dt, dt, 1e9, integrator=1, sources=[src], observers=obs)
    out = {}
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=v0, rho0=RHO0, c0=340.0, precision=prec,
                      steps_per_graph=7)
        assert dg.KERNEL_NAMES[s.info().stage_kernel] == kernel
        s.set_state(u0)
        assert s.add_monopole(-0.2, 0.1, 0.0, 0.3, 0.05, 400.0, 0.3, 1.0) > 0
        s.add_observer(*obs[0])
        r = s.run(0.0, (steps - 0.5) * dt, dt, 1e9, integrator=dg.DG_EULER)
        out[prec] = (s.get_state(), r)
        s.close()
    u, r = out[dg.DG_FP64]
    for f in range(4):
        assert rel_l2(u[f], ref["u"][f]) < 1e-10, (f, rel_l2(u[f], ref["u"][f]))
    np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(r.observers, ref["observers"], rtol=1e-8, atol=1e-14)
    assert rel_l2(out[dg.DG_MIXED][0][0], u[0]) < 1e-6


def test_graph_replay_equals_eager():
    print("case graph vs eager", flush=True)
    groups = [(ALL3, 0)]
    m, bc, gn = make_case(3, 3, 3, groups, seed=5)
    outs = []
    for use_graph in (False, True):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30, 0, 0), rho0=RHO0, c0=340.0,
                      use_graph=use_graph, steps_per_graph=7)
        s.set_state(gaussian_state(m, size=0.05))
        s.add_monopole(0.2, 0, 0, 0.3, 0.1, 300.0, 0.0, 1.0)
        r = s.run(0.0, 49.5e-5, 1e-5, 2e-4)
        outs.append((s.get_state(), r.residuals))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("kernel", ["valu", "vm", "mm", "wave"])
def test_mixed_precision_close_to_fp64(kernel, monkeypatch):
    """Mixed precision (FP32 state / fluxes, compensated RK4 accumulation) against GPU FP64,
    3D absorbing case with a monopole, 500 steps: pressure rel L2 <= 1e-6. Snapshots every
    100 steps; the mixed residuals (against the FP32 leading word of the snapshot) stay within
    1e-4 of the FP64 ones at every step."""
    print(f"case mixed vs fp64 3d kernel={kernel}", flush=True)
    monkeypatch.setenv("DGFEM_KERNEL", kernel)
    groups = [(ALL3, 0)]
    m, bc, gn = make_case(3, 6, 3, groups, jitter=0.15, seed=2)
    res = []
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30, 0, 0), rho0=RHO0, c0=340.0, precision=prec)
        assert dg.KERNEL_NAMES[s.info().stage_kernel] == kernel
        s.set_state(gaussian_state(m, size=0.02))
        s.add_monopole(0, 0, 0, 0.2, 0.1, 200.0, 0.0, 1.0)
        s.add_observer(0.5, 0.3, 0.0, 0.2)
        r = s.run(0.0, 499.5 * 2e-5, 2e-5, 100 * 2e-5)
        res.append((s.get_state(), r.observers, r.residuals))
    e = rel_l2(res[1][0][0], res[0][0][0])
    rm, rd = res[1][2][:, :5], res[0][2][:, :5]
    assert rd.shape == (500, 5) and np.all(rd[:, :4] > 0)
    er = np.abs(rm[:, :4] / rd[:, :4] - 1.0).max()
    print("mixed rel L2 p:", e, " observer:", rel_l2(res[1][1], res[0][1]), " residuals (max rel):", er)
    assert e < 1e-6
    assert rel_l2(res[1][1], res[0][1]) < 1e-6
    assert er < 1e-4


@pytest.mark.parametrize("dim,n,p,fields", [(3, 4, 3, dg.DG_FIELDS_LEE5), (3, 3, 5, dg.DG_FIELDS_LEE5),
                                            (2, 12, 4, dg.DG_FIELDS_REF4), (2, 8, 5, dg.DG_FIELDS_REF4)])
def test_mixed_precision_close_to_fp64_other_variants(dim, n, p, fields):
    """Mixed precision against GPU FP64 beyond the 4-field tetrahedra: the 5-field LEE variant
    (p = 3, 5) and triangles (p = 4, 5), absorbing walls, a monopole, 300 RK4 steps: pressure
    rel L2 <= 1e-6."""
    print(f"case mixed vs fp64 dim={dim} n={n} p={p} fields={fields}", flush=True)
    groups = [(ALL3 if dim == 3 else ALL2, 0)]
    v0 = (30.0, -10.0, 5.0) if dim == 3 else (30.0, -10.0, 0.0)
    m, bc, gn = make_case(dim, n, p, groups, jitter=0.15, seed=9)
    nf = 5 if fields == dg.DG_FIELDS_LEE5 else 4
    u0 = gaussian_state(m, size=0.02, nf=nf, c0=340.0)
    probe = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, fields=fields)
    dt = min(0.5 * probe.info().min_inradius / ((340.0 + 35.0) * (2 * p + 1)), 1e-5)
    probe.close()
    res = []
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, fields=fields,
                      precision=prec)
        s.set_state(u0)
        s.add_monopole(0.1, 0.0, 0.0, 0.2, 0.1, 300.0, 0.0, 1.0)
        s.add_observer(-0.4, 0.2, 0.0, 0.3)
        r = s.run_steps(300, dt)
        res.append((s.get_state(), r.observers))
        s.close()
    e = rel_l2(res[1][0][0], res[0][0][0])
    eo = rel_l2(res[1][1][:, :, 0], res[0][1][:, :, 0])
    print("mixed rel L2 p:", e, " observer:", eo)
    assert np.abs(res[0][0][0]).max() > 1e-3
    assert e < 1e-6 and eo < 1e-6


def test_lee5_keeps_isentropic_density():
    print("case lee5 invariant", flush=True)
    groups = [(0x10, 1), (0x2F, 0)]
    m, bc, gn = make_case(3, 3, 3, groups, seed=3)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30, 0, 0), rho0=RHO0, c0=340.0,
                  fields=dg.DG_FIELDS_LEE5)
    u0 = gaussian_state(m, 0.1, 0.2, -0.1, 0.05, nf=5, c0=340.0)
    s.set_state(u0)
    s.run(0.0, 99.5e-5 / 5, 2e-5, 1e9)
    u = s.get_state()
    assert rel_l2(u[4], u[0] / 340.0 ** 2) < 1e-12
    assert np.linalg.norm(u[0] - u0[0]) > 0


@pytest.mark.parametrize("residuals", [True, False])
@pytest.mark.parametrize("precision", [dg.DG_FP64, dg.DG_MIXED])
def test_nonfinite_is_reported(residuals, precision):
    """A run that blows up stops with DG_ERR_NONFINITE whether or not residuals are on."""
    print(f"case nonfinite residuals={residuals} precision={precision}", flush=True)
    groups = [(ALL2, 1)]
    m, bc, gn = make_case(2, 4, 2, groups)
    s = dg.Solver(m.node_coords, m.element_vertices, 2, bc, v0=(0, 0, 0), rho0=RHO0, c0=100.0,
                  precision=precision, steps_per_graph=5)
    s.set_state(gaussian_state(m, size=0.1))
    with pytest.raises(dg.DGError) as ei:
        s.run(0.0, 1e9, 10.0, 1e30, max_steps=2000, residuals=residuals)
    assert ei.value.code == dg.DG_ERR_NONFINITE
    # the failing step is the first one whose state is not finite
    n = ei.value.result.steps
    assert 1 <= n < 2000


@pytest.mark.parametrize("residuals", [True, False])
def test_nonfinite_initial_state_is_reported(residuals):
    """A non-finite value put into the state by the host is reported at the first step."""
    print(f"case nonfinite initial state residuals={residuals}", flush=True)
    groups = [(ALL2, 0)]
    m, bc, gn = make_case(2, 4, 2, groups)
    s = dg.Solver(m.node_coords, m.element_vertices, 2, bc, v0=(0, 0, 0), rho0=RHO0, c0=100.0)
    u = gaussian_state(m, size=0.1)
    u[1, 17] = np.nan
    s.set_state(u)
    with pytest.raises(dg.DGError) as ei:
        s.run(0.0, 9.5e-6, 1e-6, 1e30, residuals=residuals)
    assert ei.value.code == dg.DG_ERR_NONFINITE
    assert ei.value.result.steps == 1
