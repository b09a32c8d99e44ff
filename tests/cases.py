"""Seeded test cases shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

import dgfem_amd as dg
import oracle_py as orc

ALL2 = 0xF   # sides x-, x+, y-, y+
ALL3 = 0x3F  # plus z-, z+


def make_case(dim, n, p, groups, jitter=0.2, seed=1):
    """Jittered box mesh, boundary types by the first-node rule (library) and the group
    vertex sets the oracle resolves the same rule from."""
    m = dg.box_mesh(dim, n, p, jitter=jitter, seed=seed)
    bc = dg.resolve_bc_sides(m, groups)
    gn = orc.group_node_sets(m.element_vertices, m.face_side, groups)
    return m, bc, gn


def oracle_for(m, gn, groups, v0, rho0, c0):
    return orc.Oracle(m.node_coords, m.element_vertices, m.order, gn, [g[1] for g in groups], v0=v0, rho0=rho0,
                      c0=c0)


def smooth_state(m, seed=0, noise=0.1):
    rng = np.random.default_rng(seed)
    X = m.node_coords.reshape(-1, 3)
    u = np.stack([np.sin(2 * X[:, 0] + X[:, 1]) + rng.normal(0, noise, len(X)),
                  np.cos(X[:, 1] - X[:, 2]) + rng.normal(0, noise, len(X)),
                  0.3 * X[:, 0] * X[:, 1] + rng.normal(0, noise, len(X)),
                  (X[:, 2] ** 2 if m.dim == 3 else 0 * X[:, 2])])
    return u


def gaussian_state(m, x=0.0, y=0.0, z=0.0, size=0.1, amp=1.0, nf=4, c0=1.0):
    X = m.node_coords.reshape(-1, 3)
    u = np.zeros((nf, len(X)))
    u[0] = amp * np.exp(-((X[:, 0] - x) ** 2 + (X[:, 1] - y) ** 2 + (X[:, 2] - z) ** 2) / size)
    if nf == 5:
        u[4] = u[0] / (c0 * c0)
    return u


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
