"""CPU oracle of the learned absorbing-boundary closure (test infrastructure only: imported by
tests/ and nothing else). The closure has no counterpart in the reference solver; the
reference's characteristic absorbing flux (src/Mesh.cpp:619-666, :954-969) stays the
fallback. This module restates, in FP64 NumPy:

  * the feature map of a boundary face: 2D face frame n, t = (-n_y, n_x); per face node
    P = p/(rho0 c0), N = v.n, T = v.t, w_out = P + N, w_in = P - N; scale
    s = sqrt(mean(w_out^2 + T^2)) + 1e-30; x = [w_out/s, T/s, v0.n/c0, v0.t/c0]. 3D: two
    tangents, t1 along the tangential mean flow v0 - (v0.n) n when |v0 - (v0.n) n| > 1e-6 c0,
    else along e_a - (e_a.n) n for the axis a least aligned with n, and t2 = n x t1;
    s = sqrt(mean(w_out^2 + T1^2 + T2^2)) + 1e-30, x = [w_out/s, T1/s, T2/s, v0.n/c0,
    |v0 - (v0.n) n|/c0]. The regression target is the neighbour trace's incoming
    characteristic divided by s;
  * the network n_in -> 64 -> 64 -> n_out with tanh, evaluated with the BF16 weights
    decoded exactly to FP64, its mean-square loss and the gradient by backpropagation;
  * the exterior state that makes the Lax-Friedrichs face flux the upwind flux with
    incoming values w_hat = s y (acoustic) and 0 (vorticity).

Pins (tests/test_closure_cpu.py): the gradient against central finite differences of the
loss; BF16 rounding against torch's float32 -> bfloat16 conversion; the feature map against
rotation and scale invariance and a hand-built face; the exterior state against the upwind
flux it is built to produce.
"""
from __future__ import annotations

import numpy as np

H = 64


def bf16_to_f64(bits) -> np.ndarray:
    b = np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def f32_to_bf16(x) -> np.ndarray:
    """Round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    return ((u + 0x7FFF + lsb) >> 16).astype(np.uint16)


def sizes(nfp: int, dim: int = 2):
    return dim * nfp + 2, nfp


def nout_of(nparams: int, nin: int) -> int:
    """Output width of a network with nparams parameters and nin inputs."""
    r = nparams - (H * nin + H + H * H + H)
    assert r % (H + 1) == 0, (nparams, nin)
    return r // (H + 1)


def unpack(params, nin, nout):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    out = []
    for shape in ((H, nin), (H,), (H, H), (H,), (nout, H), (nout,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size, (o, p.size)
    return out


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([a.ravel() for a in (W1, b1, W2, b2, W3, b3)])


def forward(params, X):
    """params: FP64 values (decode BF16 bits with bf16_to_f64 first). Returns Y and the
    activations (a1, a2)."""
    X = np.asarray(X, dtype=np.float64)
    nin = X.shape[1]
    nout = nout_of(np.asarray(params).size, nin)
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, nout)
    a1 = np.tanh(X @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, (a1, a2)


def loss_grad(params, X, Y):
    """L = sum |y - Y|^2 / (n n_out) and dL/dparams (canonical order)."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n, nin = X.shape
    nout = Y.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, nout)
    y, (a1, a2) = forward(params, X)
    e = y - Y
    L = float((e * e).sum() / (n * nout))
    d3 = 2.0 * e / (n * nout)
    gW3 = d3.T @ a2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * (1.0 - a2 * a2)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (1.0 - a1 * a1)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return L, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def frame(normals, v0=(0.0, 0.0, 0.0), c0=1.0, dim=2):
    """Unit normals n [m][dim] and tangents t [m][dim-1][dim] of the faces (see above)."""
    nv = np.asarray(normals, dtype=np.float64)
    if dim == 2:
        n = nv[:, :2]
        return n, np.stack([-n[:, 1], n[:, 0]], axis=1)[:, None, :]
    n = nv[:, :3]
    v0 = np.asarray(v0, dtype=np.float64)
    vt = v0[None, :] - (n @ v0)[:, None] * n
    vtn = np.linalg.norm(vt, axis=1)
    ax = np.argmin(np.abs(n), axis=1)
    e = np.eye(3)[ax]
    alt = e - (e * n).sum(1, keepdims=True) * n
    alt /= np.linalg.norm(alt, axis=1, keepdims=True)
    use = vtn > 1e-6 * c0
    t1 = np.where(use[:, None], vt / np.where(use, vtn, 1.0)[:, None], alt)
    t2 = np.cross(n, t1)
    return n, np.stack([t1, t2], axis=1)


def characteristics(q, normals, rho0, c0, v0=(0.0, 0.0, 0.0), dim=2):
    """w_out, w_in [m][Nfp] and the tangential velocities T [m][dim-1][Nfp] of traces
    q [m][Nfp][nf]."""
    q = np.asarray(q, dtype=np.float64)
    n, t = frame(normals, v0, c0, dim)
    v = q[..., 1:1 + dim]
    P = q[..., 0] / (rho0 * c0)
    N = np.einsum("mjc,mc->mj", v, n)
    T = np.einsum("mjc,mkc->mkj", v, t)
    return P + N, P - N, T


def features(q_int, normals, v0, rho0, c0, dim=2):
    """q_int [m][Nfp][nf] interior traces -> X [m][dim Nfp + 2], s [m]."""
    v0 = np.asarray(v0, dtype=np.float64)
    wo, _, T = characteristics(q_int, normals, rho0, c0, v0, dim)
    n, t = frame(normals, v0, c0, dim)
    s = np.sqrt((wo * wo + (T * T).sum(axis=1)).mean(axis=1)) + 1e-30
    vn0 = n @ v0[:dim]
    vt0 = t[:, 0, :] @ v0[:dim] if dim == 2 else np.linalg.norm(v0[None, :] - vn0[:, None] * n, axis=1)
    parts = [wo / s[:, None]] + [T[:, k] / s[:, None] for k in range(dim - 1)]
    X = np.concatenate(parts + [vn0[:, None] / c0, vt0[:, None] / c0], axis=1)
    return X, s


def targets(q_ext, normals, rho0, c0, s, v0=(0.0, 0.0, 0.0), dim=2):
    _, wi, _ = characteristics(q_ext, normals, rho0, c0, v0, dim)
    return wi / s[:, None]


def exterior_from_incoming(q_int, normals, w_hat, v0, rho0, c0, dim=2):
    """Exterior states whose Lax-Friedrichs flux is the upwind flux with incoming acoustic
    characteristic w_hat and no incoming vorticity."""
    q = np.asarray(q_int, dtype=np.float64)
    v0 = np.asarray(v0, dtype=np.float64)
    n, t = frame(normals, v0, c0, dim)
    wo, wi, T = characteristics(q, normals, rho0, c0, v0, dim)
    vn0 = (n @ v0[:dim])[:, None]
    lam = vn0 - c0
    wip = (2 * lam * w_hat - (lam + c0) * wi) / (lam - c0)
    tp = np.where(vn0[:, None] >= 0, T, -(vn0[:, None] + c0) * T / (vn0[:, None] - c0))
    Pp, Np_ = 0.5 * (wo + wip), 0.5 * (wo - wip)
    out = q.copy()
    out[..., 0] = rho0 * c0 * Pp
    for c in range(dim):
        out[..., 1 + c] = Np_ * n[:, c:c + 1] + np.einsum("mkj,mk->mj", tp, t[:, :, c])
    if q.shape[2] == 5:
        out[..., 4] = q[..., 4] + (out[..., 0] - q[..., 0]) / (c0 * c0)
    return out


def exterior(q_int, normals, params, v0, rho0, c0, dim=2):
    """Exterior states q+ [m][Nfp][nf] the closure gives for interior traces q_int."""
    X, s = features(q_int, normals, v0, rho0, c0, dim)
    y, _ = forward(params, X)
    return exterior_from_incoming(q_int, normals, s[:, None] * y, v0, rho0, c0, dim)
