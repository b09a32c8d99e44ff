"""The split-BF16 matrix-core stage kernel of the mixed precision (DGFEM_KERNEL=bf16x6: operands
carried as 3 BF16 terms, six matrix products per contraction; bf16x3: 2 terms, three products),
tetrahedra of order 2..5, 4 and 5 fields.

Bars: one right-hand side against the oracle (the reference algorithm restated, FP64):
bf16x6 within 1e-6 relative L2 per field (FP32 operand precision; the FP32 VALU kernel of the
same precision is printed beside it), bf16x3 within 1e-4; RK4 runs in mixed precision with
bf16x6 within 1e-6 of the GPU FP64 path in pressure (field and observer series).
"""
import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL3, gaussian_state, make_case, oracle_for, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

RHO0 = 1.225


@pytest.mark.parametrize("n,p,groups,v0", [
    (2, 2, [(0x10, 1), (0x2F, 0)], (30.0, 10.0, -5.0)),
    (2, 3, [(ALL3, 0)], (30.0, 0.0, 0.0)),
    (2, 4, [(0x10, 1), (0x2F, 0)], (30.0, 0.0, 0.0)),
    (1, 5, [(ALL3, 0)], (30.0, 0.0, 0.0)),
    (2, 5, [(0x2F, 0), (0x10, 1)], (20.0, -10.0, 5.0)),
])
@pytest.mark.parametrize("kernel", ["bf16x6", "bf16x3"])
def test_bf16_rhs_matches_oracle(n, p, groups, v0, kernel, monkeypatch):
    print(f"case bf16 rhs n={n} p={p} kernel={kernel}", flush=True)
    m, bc, gn = make_case(3, n, p, groups, seed=7)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    u = smooth_state(m)
    k_ref = o.rhs(u, 1e-4)
    monkeypatch.setenv("DGFEM_KERNEL", "valu")
    sv = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, precision=dg.DG_MIXED)
    kv = sv.rhs(u, 1e-4)
    sv.close()
    monkeypatch.setenv("DGFEM_KERNEL", kernel)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, precision=dg.DG_MIXED)
    assert dg.KERNEL_NAMES[s.info().stage_kernel] == kernel
    k = s.rhs(u, 1e-4)
    bar = 1e-6 if kernel == "bf16x6" else 1e-4
    for f in range(4):
        e, ev = rel_l2(k[f], k_ref[f]), rel_l2(kv[f], k_ref[f])
        print(f"field {f}: {kernel} {e:.2e}   fp32 valu {ev:.2e}")
        assert e < bar, (f, e)


@pytest.mark.parametrize("p,n", [(3, 2), (5, 1)])
def test_bf16_lee5_rhs_matches_four_field_oracle_and_keeps_density(p, n, monkeypatch):
    print(f"case bf16 lee5 rhs p={p}", flush=True)
    groups = [(0x2F, 0), (0x10, 1)]
    v0 = (30.0, -10.0, 5.0)
    m, bc, gn = make_case(3, n, p, groups, seed=8)
    o = oracle_for(m, gn, groups, v0, RHO0, 340.0)
    u = smooth_state(m)
    u5 = np.vstack([u, u[0][None] / 340.0 ** 2])
    k_ref = o.rhs(u, 1e-4)
    monkeypatch.setenv("DGFEM_KERNEL", "bf16x6")
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, precision=dg.DG_MIXED,
                  fields=dg.DG_FIELDS_LEE5)
    k = s.rhs(u5, 1e-4)
    for f in range(4):
        assert rel_l2(k[f], k_ref[f]) < 1e-6
    assert rel_l2(k[4], k[0] / 340.0 ** 2) < 1e-6


@pytest.mark.parametrize("p,n,fields", [(3, 6, dg.DG_FIELDS_REF4), (5, 3, dg.DG_FIELDS_LEE5),
                                        (4, 3, dg.DG_FIELDS_REF4)])
def test_bf16_rk4_close_to_fp64(p, n, fields, monkeypatch):
    """Absorbing box with a monopole and an observer, 300 RK4 steps at the CFL-limited dt."""
    print(f"case bf16 rk4 vs fp64 p={p} n={n} fields={fields}", flush=True)
    groups = [(ALL3, 0)]
    v0 = (30.0, -10.0, 5.0)
    m, bc, gn = make_case(3, n, p, groups, jitter=0.15, seed=9)
    nf = 5 if fields == dg.DG_FIELDS_LEE5 else 4
    u0 = gaussian_state(m, size=0.02, nf=nf, c0=340.0)
    res = {}
    for prec, kernel in ((dg.DG_FP64, "valu"), (dg.DG_MIXED, "bf16x6"), (dg.DG_MIXED, "valu")):
        monkeypatch.setenv("DGFEM_KERNEL", kernel)
        s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, fields=fields,
                      precision=prec, steps_per_graph=10)
        dt = 0.5 * s.info().min_inradius / ((340.0 + 35.0) * (2 * p + 1))
        s.set_state(u0)
        s.add_monopole(0.1, 0.0, 0.0, 0.2, 0.1, 300.0, 0.0, 1.0)
        s.add_observer(-0.4, 0.2, 0.0, 0.3)
        r = s.run_steps(300, dt)
        res[(prec, kernel)] = (s.get_state(), r.observers)
        s.close()
    u64, o64 = res[(dg.DG_FP64, "valu")]
    for key in ((dg.DG_MIXED, "bf16x6"), (dg.DG_MIXED, "valu")):
        u, o = res[key]
        print(f"{key[1]}: pressure {rel_l2(u[0], u64[0]):.2e}, observer {rel_l2(o[:, :, 0], o64[:, :, 0]):.2e}")
    u, o = res[(dg.DG_MIXED, "bf16x6")]
    assert rel_l2(u[0], u64[0]) < 1e-6 and rel_l2(o[:, :, 0], o64[:, :, 0]) < 1e-6
    if nf == 5:
        print(f"density vs fp64 {rel_l2(u[4], u64[4]):.2e}, rho - p/c0^2 {rel_l2(u[4], u[0] / 340.0 ** 2):.2e}")
        assert rel_l2(u[4], u64[4]) < 1e-6


def test_bf16_unsupported_requests(monkeypatch):
    m, bc, _ = make_case(3, 1, 3, [(ALL3, 0)], seed=1)
    monkeypatch.setenv("DGFEM_KERNEL", "bf16x6")
    with pytest.raises(dg.DGError) as ei:
        dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(0, 0, 0), rho0=RHO0, c0=340.0, precision=dg.DG_FP64)
    assert ei.value.code == dg.DG_ERR_UNSUPPORTED
    m2, bc2, _ = make_case(2, 2, 3, [(0xF, 0)], seed=1)
    with pytest.raises(dg.DGError) as ei:
        dg.Solver(m2.node_coords, m2.element_vertices, 3, bc2, v0=(0, 0, 0), rho0=RHO0, c0=340.0,
                  precision=dg.DG_MIXED)
    assert ei.value.code == dg.DG_ERR_UNSUPPORTED


@pytest.mark.parametrize("p,fields", [(3, dg.DG_FIELDS_REF4), (5, dg.DG_FIELDS_LEE5)])
def test_bf16_blocks_with_several_groups(p, fields, monkeypatch):
    """12^3 x 6 tets (648 groups of 16 elements, more than the persistent blocks, so blocks run
    several groups and prefetch the next one), reflecting floor + absorbing, 60 RK4 steps with a
    monopole and an observer: bf16x6 within 1e-6 of FP64 in pressure (field and observer)."""
    print(f"case bf16 several groups per block p={p}", flush=True)
    groups = [(0x2F, 0), (0x10, 1)]
    v0 = (30.0, -10.0, 5.0)
    m, bc, gn = make_case(3, 12, p, groups, jitter=0.15, seed=19)
    nf = 5 if fields == dg.DG_FIELDS_LEE5 else 4
    u0 = gaussian_state(m, 0.1, -0.1, 0.2, 0.02, nf=nf, c0=340.0)
    res = {}
    for prec, kernel in ((dg.DG_FP64, "valu"), (dg.DG_MIXED, "bf16x6")):
        monkeypatch.setenv("DGFEM_KERNEL", kernel)
        s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=v0, rho0=RHO0, c0=340.0, fields=fields,
                      precision=prec, steps_per_graph=10)
        info = s.info()
        if kernel == "bf16x6":
            assert info.stage_blocks < (m.K + 15) // 16
        dt = 0.5 * info.min_inradius / ((340.0 + 35.0) * (2 * p + 1))
        s.set_state(u0)
        s.add_monopole(0.2, 0.0, 0.0, 0.1, 0.1, 300.0, 0.0, 1.0)
        s.add_observer(-0.4, 0.2, 0.0, 0.1)
        r = s.run_steps(60, dt)
        res[prec] = (s.get_state(), r.observers, r.residuals)
        s.close()
    (u64, o64, r64), (u, o, r) = res[dg.DG_FP64], res[dg.DG_MIXED]
    e, eo = rel_l2(u[0], u64[0]), rel_l2(o[:, :, 0], o64[:, :, 0])
    print(f"pressure {e:.2e}, observer {eo:.2e}")
    assert e < 1e-6 and eo < 1e-6
    assert np.abs(r[:, :4] / r64[:, :4] - 1.0).max() < 1e-4
