"""Convergence to a closed-form solution: the standing acoustic wave in a rigid box.

With no mean flow the equations are p_t + rho0 c0^2 div v = 0, v_t + grad p / rho0 = 0. In
[-1,1]^d with reflecting walls (v.n = 0) the mode P(x) = prod_i cos(pi x_i) is exact:
    p = P cos(w t),   v = -grad P sin(w t) / (rho0 w),   w = c0 pi sqrt(d).
Nodal errors at t = T on jittered meshes of n, 2n (and 4n) cells per side must fall at close to
the DG rate h^(p+1): a dropped term, a wrong sign or a wrong index in the volume, flux, lift or
reflecting-wall terms breaks the rate. The oracle (CPU) is pinned this way on small meshes; the
GPU path on finer ones, FP64 and mixed precision.
"""
import numpy as np
import pytest

import dgfem_amd as dg
from cases import ALL2, ALL3, make_case, oracle_for

RHO0, C0 = 1.3, 2.0


def exact(m, t):
    X = m.node_coords.reshape(-1, 3)
    d = m.dim
    w = C0 * np.pi * np.sqrt(d)
    c = [np.cos(np.pi * X[:, i]) for i in range(d)]
    s = [np.sin(np.pi * X[:, i]) for i in range(d)]
    P = np.prod(c, axis=0)
    u = np.zeros((4, len(X)))
    u[0] = P * np.cos(w * t)
    for i in range(d):
        dP = -np.pi * s[i] * np.prod([c[j] for j in range(d) if j != i], axis=0)
        u[1 + i] = -dP * np.sin(w * t) / (RHO0 * w)
    return u


def nodal_error(u, m, t):
    ue = exact(m, t)
    return np.linalg.norm(u[:m.dim + 1] - ue[:m.dim + 1]) / np.linalg.norm(ue[:m.dim + 1])


def steps_for(n, p, T, cfl=0.1):
    """Steps of a dt well inside the RK4 limit (time error far below the spatial error)."""
    h = 2.0 / n
    return int(np.ceil(T / (cfl * h / (C0 * (2 * p + 1)))))


def run_oracle(dim, n, p, T):
    groups = [(ALL2 if dim == 2 else ALL3, 1)]
    m, bc, gn = make_case(dim, n, p, groups, jitter=0.1, seed=11)
    o = oracle_for(m, gn, groups, (0.0, 0.0, 0.0), RHO0, C0)
    k = steps_for(n, p, T)
    dt = T / k
    r = o.run(exact(m, 0.0), 0.0, (k - 0.5) * dt, dt, 1e9)
    assert r["steps"] == k
    return nodal_error(r["u"], m, k * dt)


def rates(errs):
    return [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


@pytest.mark.parametrize("dim,p,ns", [(2, 1, (4, 8, 16)), (2, 2, (4, 8, 16)), (2, 3, (4, 8)), (3, 1, (4, 8)),
                                      (3, 2, (4, 8))])
def test_oracle_converges_at_the_dg_rate(dim, p, ns):
    # the 3D meshes affordable here (8 cells per side at most) are still pre-asymptotic
    T = 0.25
    errs = [run_oracle(dim, n, p, T) for n in ns]
    r = rates(errs)
    print(dim, p, errs, r)
    assert errs[-1] < (0.05 if dim == 2 else 0.2)
    assert min(r) > p + 1 - (0.6 if dim == 2 else 0.8), (errs, r)


def run_gpu(dim, n, p, T, precision):
    groups = [(ALL2 if dim == 2 else ALL3, 1)]
    m, bc, _ = make_case(dim, n, p, groups, jitter=0.1, seed=11)
    s = dg.Solver(m.node_coords, m.element_vertices, p, bc, v0=(0.0, 0.0, 0.0), rho0=RHO0, c0=C0,
                  precision=precision)
    k = steps_for(n, p, T)
    dt = T / k
    s.set_state(exact(m, 0.0))
    s.run_steps(k, dt)
    e = nodal_error(s.get_state(), m, k * dt)
    s.close()
    return e


@pytest.mark.gpu
@pytest.mark.parametrize("dim,p,ns", [(2, 2, (8, 16, 32)), (2, 4, (4, 8, 16)), (3, 2, (4, 8, 16)), (3, 3, (4, 8))])
def test_gpu_converges_at_the_dg_rate(dim, p, ns):
    print(f"case convergence dim={dim} p={p} n={ns}", flush=True)
    T = 0.25
    for prec in (dg.DG_FP64, dg.DG_MIXED):
        errs = [run_gpu(dim, n, p, T, prec) for n in ns]
        r = rates(errs)
        print(prec, errs, r, flush=True)
        assert r[-1] > p + 1 - 0.5 and min(r) > p + 1 - 1.0, (prec, errs, r)
