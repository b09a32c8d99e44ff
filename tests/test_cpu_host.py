"""CPU tests of the host logic: the library's FP64 precompute and nodal operators against
the oracle (the reference's quadrature algorithm), and the C ABI surface."""
import os
import re

import numpy as np
import pytest

import dgfem_amd as dg
import nodal_ref
from cases import ALL2, ALL3, make_case, oracle_for, rel_l2, smooth_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dim,n,p,groups,v0", [
    (2, 4, 1, [(ALL2, 1)], (-30, 30, 0)),
    (2, 4, 3, [(ALL2, 1)], (-30, 30, 0)),
    (2, 3, 5, [(0x3, 1), (0xC, 0)], (10, 20, 0)),
    (3, 2, 2, [(0x10, 1), (0x2F, 0)], (30, 10, -5)),
    (3, 2, 3, [(ALL3, 0)], (30, 0, 0)),
    (3, 1, 5, [(ALL3, 0)], (30, 0, 0)),
])
def test_nodal_precompute_matches_oracle(dim, n, p, groups, v0):
    m, bc, gn = make_case(dim, n, p, groups, seed=3)
    o = oracle_for(m, gn, groups, v0, 1.225, 100.0)
    pre = dg.mesh_precompute(m.node_coords, m.element_vertices, p, bc)
    ops = dg.ref_operators(dim, p)
    u = smooth_state(m)
    ko = o.rhs(u, 1e-3)
    kn = nodal_ref.nodal_rhs(u, pre, ops, dim, v0, 1.225, 100.0, 1e-3)
    for f in range(4):
        if np.linalg.norm(ko[f]) > 0:
            assert rel_l2(kn[f], ko[f]) < 1e-12


def test_abi_exports_every_declared_symbol():
    lib = dg.load_library()
    hdr = open(os.path.join(ROOT, "include", "dgfem_amd.h")).read()
    names = re.findall(r"DG_API\s+[\w\s\*]+?\b(dg_\w+)\s*\(", hdr)
    assert len(names) >= 20
    for nm in names:
        assert hasattr(lib, nm), nm


def test_box_mesh_generator_invariants():
    for dim, n in ((2, 6), (3, 3)):
        m = dg.box_mesh(dim, n, 2, jitter=0.2, seed=11)
        pre = dg.mesh_precompute(m.node_coords, m.element_vertices, 2)
        # total measure of the box and positive orientation
        detJ = 1.0 / np.linalg.det(pre["Jinv"][:, :dim, :dim])
        assert (detJ > 0).all()
        fact = 2.0 if dim == 2 else 6.0
        assert abs(detJ.sum() / fact - 2.0 ** dim) < 1e-12
        # every interior face pairs two elements reciprocally; boundary count matches
        nb = pre["nbr"]
        nbound = (nb < 0).sum()
        assert nbound == (4 * n if dim == 2 else 12 * n * n)
        assert ((m.face_side >= 0) == (nb < 0)).all()
        K = m.K
        for e in range(K):
            for f in range(dim + 1):
                if nb[e, f] >= 0:
                    fp = pre["code"][e, f] & 3
                    assert nb[nb[e, f], fp] == e
