"""Behaviour of the C ABI across calls, on the GPU: successive runs on one context (changed dt,
changed integrator) continue from the state the previous run left, as the oracle does for the
same sequence; state round trips; status codes of invalid calls and of a failing callback."""
import ctypes as C

import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL2, gaussian_state, make_case, oracle_for, rel_l2

pytestmark = pytest.mark.gpu

RHO0, C0 = 1.225, 100.0
V0 = (-30.0, 20.0, 0.0)


def _case():
    groups = [(0x3, 1), (0xC, 0)]
    m, bc, gn = make_case(2, 6, 3, groups, jitter=0.15, seed=12)
    return m, bc, oracle_for(m, gn, groups, V0, RHO0, C0)


@pytest.mark.parametrize("use_graph", [True, False])
def test_successive_runs_continue_the_state(use_graph):
    print(f"case successive runs graph={use_graph}", flush=True)
    m, bc, o = _case()
    u = gaussian_state(m, 0.1, -0.1, 0.0, 0.05)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0, use_graph=use_graph,
                  steps_per_graph=6)
    s.set_state(u)
    t = 0.0
    for dt, steps, integ in ((2e-5, 30, dg.DG_RK4), (1.3e-5, 17, dg.DG_RK4), (4e-6, 9, dg.DG_EULER),
                             (2e-5, 11, dg.DG_RK4)):
        t_end = t + (steps - 0.5) * dt
        ref = o.run(u, t, t_end, dt, 1e9, integrator=0 if integ == dg.DG_RK4 else 1)
        r = s.run(t, t_end, dt, 1e9, integrator=integ)
        assert r.steps == ref["steps"] == steps
        np.testing.assert_allclose(r.residuals, ref["residuals"], rtol=1e-8, atol=1e-300)
        got = s.get_state()
        for f in range(3):
            assert rel_l2(got[f], ref["u"][f]) < 1e-10, (dt, integ, f, rel_l2(got[f], ref["u"][f]))
        u = ref["u"]
        t = t + steps * dt
    s.close()


def test_state_round_trips():
    print("case state round trip", flush=True)
    m, bc, _ = _case()
    rng = np.random.default_rng(4)
    u = rng.normal(size=(4, m.K * m.Np)) * np.exp(rng.uniform(-20, 20, size=(4, m.K * m.Np)))
    s64 = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0)
    s64.set_state(u)
    assert np.array_equal(s64.get_state(), u)
    smx = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0, precision=dg.DG_MIXED)
    smx.set_state(u)
    # compensated pair: the FP32 leading word plus the FP32 rounding error of it
    assert np.all(np.abs(smx.get_state() - u) <= 2.0 ** -46 * np.abs(u))
    s64.close()
    smx.close()


def test_zero_steps_and_invalid_calls():
    print("case invalid calls", flush=True)
    m, bc, _ = _case()
    u = gaussian_state(m)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0)
    s.set_state(u)
    r = s.run(1.0, 0.5, 1e-5, 1e9)
    assert r.steps == 0 and np.array_equal(s.get_state(), u)
    for dt in (0.0, -1e-5, float("nan")):
        with pytest.raises(dg.DGError) as e:
            s.run(0.0, 1e-3, dt, 1e9)
        assert e.value.code == dg.DG_ERR_INVALID
    with pytest.raises(dg.DGError) as e:
        s.run(0.0, 1e-3, 1e-5, 1e9, integrator=7)
    assert e.value.code == dg.DG_ERR_INVALID
    with pytest.raises(dg.DGError) as e:
        s.set_closure(np.zeros(dg.closure_num_params(4), dtype=np.uint16))
    assert e.value.code == dg.DG_ERR_INVALID
    with pytest.raises(dg.DGError) as e:
        s.face_traces(np.array([m.K], dtype=np.int64), np.array([0], dtype=np.int32))
    assert e.value.code == dg.DG_ERR_INVALID
    assert np.array_equal(s.get_state(), u)
    s.close()


def test_failing_step_callback_stops_the_run():
    print("case failing callback", flush=True)
    m, bc, _ = _case()
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0, steps_per_graph=5)
    s.set_state(gaussian_state(m))
    seen = []

    def step_cb(_user, recp):
        seen.append(recp.contents.step)
        return 1 if recp.contents.step == 7 else 0

    def save_cb(_user, step, t, state):
        return 0

    prm = dg.RunParams(0.0, 1.0, 1e-5, 1e9, dg.DG_RK4, 1, -1)
    done = C.c_int64()
    st = s._lib.dg_run(s._h, C.byref(prm), dg.STEP_FN(step_cb), dg.SAVE_FN(save_cb), None, C.byref(done))
    assert st == dg.DG_ERR_CALLBACK
    assert done.value == 8 and seen == list(range(8))
    assert "callback" in s._lib.dg_last_error(s._h).decode()
    s.close()


def test_independent_contexts_interleaved_and_concurrent():
    """Contexts share nothing: two solvers on different meshes and precisions, run interleaved
    and then from two host threads at once, end in the same states as when run alone."""
    import threading
    print("case two contexts", flush=True)
    m1, bc1, _ = _case()
    m2, bc2, _ = make_case(3, 3, 2, [(0x3F, 0)], jitter=0.15, seed=13)
    u1, u2 = gaussian_state(m1, size=0.05), gaussian_state(m2, size=0.05)

    def solver(m, bc, p, prec):
        s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=V0, rho0=RHO0, c0=C0, precision=prec,
                      steps_per_graph=4)
        return s

    alone = []
    for m, bc, p, prec, u in ((m1, bc1, 3, dg.DG_FP64, u1), (m2, bc2, 2, dg.DG_MIXED, u2)):
        s = solver(m, bc, p, prec)
        s.set_state(u)
        s.run_steps(40, 1e-5)
        alone.append(s.get_state())
        s.close()
    a, b = solver(m1, bc1, 3, dg.DG_FP64), solver(m2, bc2, 2, dg.DG_MIXED)
    a.set_state(u1)
    b.set_state(u2)
    for _ in range(4):
        a.run_steps(10, 1e-5)
        b.run_steps(10, 1e-5)
    assert np.array_equal(a.get_state(), alone[0]) and np.array_equal(b.get_state(), alone[1])
    a.set_state(u1)
    b.set_state(u2)
    errors = []

    def work(s):
        try:
            s.run_steps(40, 1e-5)
        except Exception as ex:  # reported below
            errors.append(ex)

    th = [threading.Thread(target=work, args=(s,)) for s in (a, b)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors
    assert np.array_equal(a.get_state(), alone[0]) and np.array_equal(b.get_state(), alone[1])
    a.close()
    b.close()


def test_residuals_of_a_large_grid_match_the_host_sum():
    """More than 16,384 stage blocks (56^3 x 6 linear tets, 64 per block): the residual partial
    sums are folded in groups before the end-of-step kernel. The last step's residuals against
    the snapshot taken at step 0 equal the mean squared differences summed on the host from the
    returned states."""
    print("case residuals large grid 56^3x6 p1", flush=True)
    m = dg.box_mesh(3, 56, 1, jitter=0.1, seed=14)
    bc = dg.resolve_bc_sides(m, [(0x3F, dg.DG_BC_ABSORBING)])
    c0 = 340.0
    s = dg.Solver(m.node_coords, m.element_vertices, 1, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=c0)
    assert m.K // 64 > 16384
    u0 = gaussian_state(m, 0.1, 0.0, -0.1, 0.05)
    s.set_state(u0)
    r = s.run_steps(3, 2e-6)
    u = s.get_state()
    d = u - u0
    host = [np.mean(d[0] ** 2), np.mean((d[0] / c0 ** 2) ** 2), np.mean(d[1] ** 2), np.mean(d[2] ** 2),
            np.mean(d[3] ** 2)]
    np.testing.assert_allclose(r.residuals[-1], host, rtol=1e-10, atol=0)
    assert min(host) > 0
    s.close()


@pytest.mark.parametrize("overlap", [False, True])
@pytest.mark.parametrize("precision", [dg.DG_FP64, dg.DG_MIXED])
def test_captured_steps_match_eager_steps(precision, overlap, monkeypatch):
    """Captured RK4 sequences may run the end-of-step kernel on a branch of its own and the next
    step's source overwrite on the main branch when no observer node is a source node (else the
    single end-of-step kernel, as the eager path always does); DGFEM_POST_BRANCH=1 asks for the
    branch on this small grid. Both must give the eager results bit for bit: state, residuals
    and observers over batches of several sequence lengths, with a monopole that switches off
    mid-run and a table source."""
    print(f"case captured vs eager precision={precision} overlap={overlap}", flush=True)
    monkeypatch.setenv("DGFEM_POST_BRANCH", "1")
    m, bc, _ = _case()
    u = gaussian_state(m, -0.3, 0.2, 0.0, 0.05)
    out = []
    for use_graph in (False, True):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=V0, rho0=RHO0, c0=C0, precision=precision,
                      use_graph=use_graph, steps_per_graph=4)
        s.add_monopole(0.1, -0.1, 0.0, 0.3, 0.5, 900.0, 0.2, 31 * 2e-5)
        s.add_table_source(-0.5, -0.5, 0.0, 0.25, lambda t: 0.2 * np.cos(3000.0 * t) if t < 40 * 2e-5 else None)
        s.add_observer(0.6, 0.5, 0.0, 0.3)
        s.add_observer(-0.6, 0.4, 0.0, 0.2)
        if overlap:
            s.add_observer(0.1, -0.1, 0.0, 0.2)
        assert s.info().post_branch == (1 if use_graph and not overlap else 0)
        s.set_state(u)
        r = s.run(0.0, 52.5 * 2e-5, 2e-5, 9 * 2e-5)
        assert r.steps == 53
        out.append((s.get_state(), r.residuals, r.observers))
        s.close()
    (u0, r0, o0), (u1, r1, o1) = out
    np.testing.assert_array_equal(u1, u0)
    np.testing.assert_array_equal(r1, r0)
    np.testing.assert_array_equal(o1, o0)
    assert np.abs(o0[:, 0]).max() > 0


@pytest.fixture(scope="module")
def large_grid():
    # 41^3 x 6 p=3 tets: 16,541 stage blocks of 25 elements (more than 4 x 4,096, so the residual
    # partials are folded in groups, and more than 4,096, so the default rule takes the branch)
    m = dg.box_mesh(3, 41, 3, jitter=0.15, seed=21)
    bc = dg.resolve_bc_sides(m, [(0x3F, dg.DG_BC_ABSORBING)])
    return m, bc


@pytest.mark.parametrize("precision", [dg.DG_FP64, dg.DG_MIXED])
def test_captured_steps_match_eager_on_a_large_grid(large_grid, precision, monkeypatch):
    """The default branch rule on a grid of more than 16,384 stage blocks: captured sequences run
    the end-of-step kernel on its branch (asserted through dg_get_info) with a monopole and two
    observers, grouped residual sums and a save step inside the run; state, residuals and
    observers equal the eager run's bit for bit."""
    print(f"case captured vs eager large grid precision={precision}", flush=True)
    monkeypatch.delenv("DGFEM_POST_BRANCH", raising=False)
    m, bc = large_grid
    u = gaussian_state(m, 0.1, -0.2, 0.05, 0.02)
    out = []
    for use_graph in (False, True):
        s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=RHO0, c0=340.0,
                      precision=precision, use_graph=use_graph, steps_per_graph=4)
        s.add_monopole(0.0, 0.0, 0.0, 0.1, 0.1, 2000.0, 0.0, 1.0)
        s.add_observer(0.5, 0.0, 0.0, 0.1)
        s.add_observer(0.0, -0.5, 0.25, 0.1)
        info = s.info()
        # 25 elements per 512-thread block (VALU kernel), 24 with the volume on the matrix cores
        eb = 24 if dg.KERNEL_NAMES[info.stage_kernel] == "vm" else 25
        assert info.stage_blocks > 16384 and info.elements_per_block == eb
        assert info.post_branch == (1 if use_graph else 0)
        s.set_state(u)
        r = s.run(0.0, 10.5 * 2e-6, 2e-6, 6 * 2e-6)
        assert r.steps == 11
        out.append((s.get_state(), r.residuals, r.observers))
        s.close()
    (u0, r0, o0), (u1, r1, o1) = out
    np.testing.assert_array_equal(u1, u0)
    np.testing.assert_array_equal(r1, r0)
    np.testing.assert_array_equal(o1, o0)
    assert np.abs(o0[:, :, 0]).max() > 0 and r0[:, 0].min() > 0
