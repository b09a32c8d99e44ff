"""NumPy restatement of the device kernel's nodal right-hand side (test infrastructure).

Uses the host precompute exported by the library (operators, geometry, neighbours) and
evaluates rhs = sum_x A_x sum_r (dxi_r/dx) Dw_r q - sum_f Fscale_f LIFT_f pi_f exactly as
dg_kernels.hpp does, in FP64 on the CPU. Tests compare it with the oracle (the reference
algorithm) to pin the host precompute without a GPU.
"""
from __future__ import annotations

import numpy as np


def nodal_rhs(u, pre, ops, dim, v0, rho0, c0, dt, nf=4, closure=None):
    """closure(q_int [m][Nfp][nf], normals [m][3]) -> exterior states of the learned
    faces (neighbour code <= -3), used in the interior flux formula."""
    tc = pre["to_canon"]
    K, Np = tc.shape
    N = K * Np
    u = np.asarray(u, dtype=np.float64).reshape(nf, N)
    q = np.zeros((K, Np, nf))
    rows = np.arange(K)[:, None]
    q[rows, tc, :] = u.T.reshape(K, Np, nf)
    Ji = pre["Jinv"]
    Dw, L, fm, Nref = ops["Dw"], ops["LIFT"], ops["fmask"], ops["Nref"]
    Nfp = ops["Nfp"]
    Dq = [np.einsum("ij,ejf->eif", Dw[r], q) for r in range(dim)]
    W = [sum(Ji[:, r, x][:, None, None] * Dq[r] for r in range(dim)) for x in range(dim)]
    rhs = np.zeros_like(q)
    rc2 = rho0 * c0 * c0
    for x in range(dim):
        rhs[..., 0] += v0[x] * W[x][..., 0] + rc2 * W[x][..., 1 + x]
        for kv in range(3):
            rhs[..., 1 + kv] += v0[x] * W[x][..., 1 + kv]
        if nf == 5:
            rhs[..., 4] += v0[x] * W[x][..., 4] + rho0 * W[x][..., 1 + x]
    for kv in range(dim):
        rhs[..., 1 + kv] += W[kv][..., 0] / rho0
    nbr, code, fperm = pre["nbr"], pre["code"], pre["face_perm"]
    v0 = np.asarray(v0, dtype=np.float64)
    for f in range(dim + 1):
        nv = np.zeros((K, 3))
        for c in range(dim):
            nv[:, c] = sum(Ji[:, r, c] * Nref[f, r] for r in range(dim))
        fs = np.linalg.norm(nv, axis=1)
        n = nv / fs[:, None]
        qm = q[:, fm[f], :]                      # [K, Nfp, nf]
        nb = nbr[:, f]
        qp = np.zeros_like(qm)
        inner = nb >= 0
        fp = (code[:, f] & 3).astype(int)
        pc = (code[:, f] >> 2).astype(int)
        idx = fm[fp[inner]][np.arange(inner.sum())[:, None], fperm[pc[inner]]]
        qp[inner] = q[nb[inner][:, None], idx, :]
        vn0 = n @ v0

        def fnorm(qq):
            nvq = np.einsum("kc,kac->ka", n, qq[..., 1:4])
            out = np.zeros_like(qq)
            out[..., 0] = vn0[:, None] * qq[..., 0] + rc2 * nvq
            for r in range(3):
                out[..., 1 + r] = vn0[:, None] * qq[..., 1 + r] + n[:, r][:, None] * qq[..., 0] / rho0
            if nf == 5:
                out[..., 4] = vn0[:, None] * qq[..., 4] + rho0 * nvq
            return out

        learned = nb <= -3
        if learned.any():
            qp[learned] = closure(qm[learned], n[learned])
        inner = inner | learned
        pi = np.zeros_like(qm)
        # interior
        pi[inner] = (0.5 * (fnorm(qm) + fnorm(qp)) + 0.5 * c0 * (qm - qp))[inner]
        # reflecting
        refl = nb == -2
        qr = qm.copy()
        vn = np.einsum("kc,kac->ka", n, qm[..., 1:4])
        qr[..., 1:4] -= vn[..., None] * n[:, None, :]
        pi[refl] = fnorm(qr)[refl]
        # absorbing
        absb = nb == -1
        a = c0 + vn0
        lam = np.where(vn0 < 0, 0.0, -1.0)
        nvm = vn
        pa = np.zeros_like(qm)
        pa[..., 0] = 0.25 * (a[:, None] * qm[..., 0] + c0 * rho0 * a[:, None] * nvm)
        for r in range(3):
            tang = (qm[..., 1 + r] - n[:, r][:, None] * nvm) if (dim == 3 or r < 2) else 0.0 * nvm
            pa[..., 1 + r] = 0.25 * (n[:, r][:, None] * a[:, None] * qm[..., 0] / (rho0 * c0)
                                     + a[:, None] * n[:, r][:, None] * nvm - (vn0 * lam)[:, None] * tang)
        if nf == 5:
            pa[..., 4] = 0.25 * (a[:, None] * qm[..., 4] + rho0 / c0 * a[:, None] * nvm)
        pi[absb] = pa[absb]
        rhs -= np.einsum("ia,kaf->kif", L[:, f * Nfp:(f + 1) * Nfp], fs[:, None, None] * pi)
    k = dt * rhs
    out = np.zeros((nf, N))
    out[:, :] = k[rows, tc, :].reshape(N, nf).T
    return out
