"""ctypes binding of libdgfem_amd.so (the C ABI declared in include/dgfem_amd.h).

Python mirror of the reference's time-integration interface for the hot path:
``solver::rungeKutta(u, mesh, config)`` / ``solver::forwardEuler`` (include/solver.h:16-25,
src/solver.cpp:66-584). The compute runs in the HIP library; this module only moves
arrays across the boundary and raises when the library or the device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DGFEM_LIB selects another build of the library (e.g. a variant kept beside the default).
LIB_PATH = os.environ.get("DGFEM_LIB") or os.path.join(_HERE, "libdgfem_amd.so")

DG_OK = 0
DG_ERR_INVALID, DG_ERR_MESH, DG_ERR_DEVICE, DG_ERR_NONFINITE, DG_ERR_UNSUPPORTED, DG_ERR_CALLBACK = 1, 2, 3, 4, 5, 6
DG_FP64, DG_MIXED = 0, 1
DG_FIELDS_REF4, DG_FIELDS_LEE5 = 0, 1
DG_BC_ABSORBING, DG_BC_REFLECTING, DG_BC_LEARNED = 0, 1, 2
DG_RK4, DG_EULER = 0, 1
DG_KERNEL_VALU, DG_KERNEL_MM, DG_KERNEL_WAVE, DG_KERNEL_BF16X6, DG_KERNEL_BF16X3, DG_KERNEL_VM = 0, 1, 2, 3, 4, 5
KERNEL_NAMES = {DG_KERNEL_VALU: "valu", DG_KERNEL_MM: "mm", DG_KERNEL_WAVE: "wave", DG_KERNEL_BF16X6: "bf16x6",
                DG_KERNEL_BF16X3: "bf16x3", DG_KERNEL_VM: "vm"}


class DGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dg status {code}] {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("order", C.c_int32), ("num_elements", C.c_int64),
                ("nodes_per_element", C.c_int32), ("node_coords", C.POINTER(C.c_double)),
                ("element_vertices", C.POINTER(C.c_int64)), ("face_bc", C.POINTER(C.c_int8))]


class Physics(C.Structure):
    _fields_ = [("v0", C.c_double * 3), ("rho0", C.c_double), ("c0", C.c_double), ("fields", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("use_graph", C.c_int32), ("steps_per_graph", C.c_int32),
                ("device", C.c_int32)]


class RunParams(C.Structure):
    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("dt", C.c_double), ("t_rate", C.c_double),
                ("integrator", C.c_int32), ("compute_residuals", C.c_int32), ("max_steps", C.c_int64)]


class StepRecord(C.Structure):
    _fields_ = [("step", C.c_int64), ("t", C.c_double), ("residuals", C.POINTER(C.c_double)),
                ("num_observers", C.c_int32), ("observers", C.POINTER(C.c_double)), ("step_seconds", C.c_double)]


class TrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_double), ("beta1", C.c_double),
                ("beta2", C.c_double), ("eps", C.c_double), ("seed", C.c_uint64)]


class Info(C.Structure):
    _fields_ = [("dim", C.c_int32), ("order", C.c_int32), ("nodes_per_element", C.c_int32),
                ("nodes_per_face", C.c_int32), ("fields", C.c_int32), ("precision", C.c_int32),
                ("num_elements", C.c_int64), ("num_faces", C.c_int64), ("num_boundary_faces", C.c_int64),
                ("min_inradius", C.c_double), ("device_bytes", C.c_int64), ("stage_kernel", C.c_int32),
                ("post_branch", C.c_int32), ("elements_per_block", C.c_int32), ("stage_blocks", C.c_int64)]


VALUE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_double))
STEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(StepRecord))
SAVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_double))

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C dgfem-acoustic_amd` "
                           "(the solver has no CPU fallback)")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "dg_create": (C.c_int32, [P(MeshDesc), P(Physics), P(Options), P(C.c_void_p)]),
        "dg_destroy": (None, [C.c_void_p]),
        "dg_last_error": (C.c_char_p, [C.c_void_p]),
        "dg_set_state": (C.c_int32, [C.c_void_p, P(C.c_double)]),
        "dg_get_state": (C.c_int32, [C.c_void_p, P(C.c_double)]),
        "dg_add_gaussian": (C.c_int32, [C.c_void_p] + [C.c_double] * 5),
        "dg_add_source_monopole": (C.c_int32, [C.c_void_p] + [C.c_double] * 8 + [P(C.c_int64)]),
        "dg_add_source_table": (C.c_int32, [C.c_void_p] + [C.c_double] * 4 + [VALUE_FN, C.c_void_p, P(C.c_int64)]),
        "dg_add_observer": (C.c_int32, [C.c_void_p] + [C.c_double] * 4 + [P(C.c_int64)]),
        "dg_run": (C.c_int32, [C.c_void_p, P(RunParams), STEP_FN, SAVE_FN, C.c_void_p, P(C.c_int64)]),
        "dg_run_collect": (C.c_int32, [C.c_void_p, P(RunParams), P(C.c_double), P(C.c_double), P(C.c_double),
                                       C.c_int64, P(C.c_int64)]),
        "dg_rhs": (C.c_int32, [C.c_void_p, P(C.c_double), P(C.c_double), C.c_double]),
        "dg_get_info": (C.c_int32, [C.c_void_p, P(Info)]),
        "dg_profile_steps": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, P(C.c_double), P(C.c_double),
                                         P(C.c_double)]),
        "dg_last_run_device_ms": (C.c_double, [C.c_void_p]),
        "dg_box_mesh_size": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, P(C.c_int64), P(C.c_int32)]),
        "dg_box_mesh_fill": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_double,
                                         C.c_double, P(C.c_double), P(C.c_int64), P(C.c_int8)]),
        "dg_resolve_bc_sides": (C.c_int32, [C.c_int32, C.c_int64, P(C.c_int64), P(C.c_int8), C.c_int32,
                                            P(C.c_uint32), P(C.c_int8), P(C.c_int8)]),
        "dg_ref_operators": (C.c_int32, [C.c_int32, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_int32),
                                         P(C.c_double), P(C.c_int32)]),
        "dg_eval_formula": (C.c_double, [C.c_char_p, C.c_double]),
        "dg_read_source_table": (C.c_int32, [C.c_char_p, P(C.c_double), C.c_int64, P(C.c_int64)]),
        "dg_case_summary": (C.c_int32, [C.c_char_p, C.c_char_p, C.c_int64]),
        "dg_ref_mass": (C.c_int32, [C.c_int32, C.c_int32, P(C.c_double)]),
        "dg_mesh_precompute": (C.c_int32, [P(MeshDesc), P(C.c_double), P(C.c_int32), P(C.c_uint8), P(C.c_int32),
                                           P(C.c_int32)]),
        "dg_set_closure": (C.c_int32, [C.c_void_p, P(C.c_uint16), C.c_int64]),
        "dg_learned_faces": (C.c_int32, [C.c_void_p, P(C.c_int64), P(C.c_int32), C.c_int64, P(C.c_int64)]),
        "dg_face_traces": (C.c_int32, [C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_int32), P(C.c_double),
                                       P(C.c_double)]),
        "dg_face_traces_noalloc": (C.c_int32, [C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_int32), P(C.c_double),
                                               P(C.c_double)]),
        "dg_closure_num_params": (C.c_int64, [C.c_int32, C.c_int32]),
        "dg_closure_last_error": (C.c_char_p, []),
        "dg_closure_to_bf16": (None, [P(C.c_float), C.c_int64, P(C.c_uint16)]),
        "dg_closure_features": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int64, P(C.c_double), P(C.c_double),
                                            P(Physics), P(C.c_float), P(C.c_float), P(C.c_float)]),
        "dg_closure_forward": (C.c_int32, [C.c_int32, C.c_int32, P(C.c_uint16), P(C.c_float), C.c_int64,
                                           P(C.c_float), C.c_int32]),
        "dg_closure_loss_grad": (C.c_int32, [C.c_int32, C.c_int32, P(C.c_uint16), P(C.c_float), P(C.c_float),
                                             C.c_int64, P(C.c_double), P(C.c_float), C.c_int32]),
        "dg_closure_train": (C.c_int32, [C.c_int32, C.c_int32, P(C.c_float), P(C.c_uint16), P(C.c_float),
                                         P(C.c_float), C.c_int64, P(TrainParams), P(C.c_double), C.c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class BoxMesh:
    dim: int
    order: int
    n: int
    node_coords: np.ndarray     # [K, Np, 3]
    element_vertices: np.ndarray  # [K, dim+1]
    face_side: np.ndarray       # [K, dim+1]

    @property
    def K(self) -> int:
        return self.node_coords.shape[0]

    @property
    def Np(self) -> int:
        return self.node_coords.shape[1]


def box_mesh(dim: int, n: int, order: int, jitter: float = 0.0, seed: int = 1, lo: float = -1.0,
             hi: float = 1.0) -> BoxMesh:
    """Seeded jittered box mesh (triangles split along random diagonals / Kuhn tetrahedra)."""
    lib = load_library()
    K = C.c_int64()
    Np = C.c_int32()
    st = lib.dg_box_mesh_size(dim, n, order, C.byref(K), C.byref(Np))
    if st != DG_OK:
        raise DGError(st, "dg_box_mesh_size failed")
    X = np.zeros((K.value, Np.value, 3))
    V = np.zeros((K.value, dim + 1), dtype=np.int64)
    S = np.zeros((K.value, dim + 1), dtype=np.int8)
    st = lib.dg_box_mesh_fill(dim, n, order, jitter, seed, lo, hi, _ptr(X, C.c_double), _ptr(V, C.c_int64),
                              _ptr(S, C.c_int8))
    if st != DG_OK:
        raise DGError(st, "dg_box_mesh_fill failed")
    return BoxMesh(dim, order, n, X, V, S)


def resolve_bc_sides(mesh: BoxMesh, groups: list[tuple[int, int]]) -> np.ndarray:
    """groups: [(side_mask, bc_type)] in ascending physical-tag order -> face_bc [K, dim+1]."""
    lib = load_library()
    masks = np.array([g[0] for g in groups], dtype=np.uint32)
    types = np.array([g[1] for g in groups], dtype=np.int8)
    out = np.zeros_like(mesh.face_side)
    st = lib.dg_resolve_bc_sides(mesh.dim, mesh.K, _ptr(mesh.element_vertices, C.c_int64),
                                 _ptr(mesh.face_side, C.c_int8), len(groups), _ptr(masks, C.c_uint32),
                                 _ptr(types, C.c_int8), _ptr(out, C.c_int8))
    if st != DG_OK:
        raise DGError(st, "dg_resolve_bc_sides failed")
    return out


def eval_formula(expr: str, t: float) -> float:
    return load_library().dg_eval_formula(expr.encode(), t)


def read_source_table(path: str) -> np.ndarray:
    lib = load_library()
    n = C.c_int64()
    st = lib.dg_read_source_table(path.encode(), None, 0, C.byref(n))
    if st != DG_OK:
        raise DGError(st, f"cannot read source table {path}")
    rows = np.zeros((n.value, 2))
    lib.dg_read_source_table(path.encode(), _ptr(rows, C.c_double), n.value, C.byref(n))
    return rows


def case_summary(config_path: str) -> dict:
    import json
    lib = load_library()
    buf = C.create_string_buffer(1 << 20)
    st = lib.dg_case_summary(config_path.encode(), buf, len(buf))
    if st != DG_OK:
        raise DGError(st, buf.value.decode())
    return json.loads(buf.value.decode())


def num_nodes(dim: int, p: int) -> int:
    return (p + 1) * (p + 2) // 2 if dim == 2 else (p + 1) * (p + 2) * (p + 3) // 6


def ref_operators(dim: int, p: int) -> dict:
    """Reference-element operators of the host precompute (canonical lattice order)."""
    lib = load_library()
    Np = num_nodes(dim, p)
    Nfp = p + 1 if dim == 2 else num_nodes(2, p)
    Dw = np.zeros((dim, Np, Np))
    L = np.zeros((Np, (dim + 1) * Nfp))
    fm = np.zeros((dim + 1, Nfp), dtype=np.int32)
    Nr = np.zeros((dim + 1, 3))
    lat = np.zeros((Np, 4), dtype=np.int32)
    st = lib.dg_ref_operators(dim, p, _ptr(Dw, C.c_double), _ptr(L, C.c_double), _ptr(fm, C.c_int32),
                              _ptr(Nr, C.c_double), _ptr(lat, C.c_int32))
    if st != DG_OK:
        raise DGError(st, "dg_ref_operators failed")
    M = np.zeros((Np, Np))
    st = lib.dg_ref_mass(dim, p, _ptr(M, C.c_double))
    if st != DG_OK:
        raise DGError(st, "dg_ref_mass failed")
    return dict(Dw=Dw, LIFT=L, fmask=fm, Nref=Nr, lattice=lat, Np=Np, Nfp=Nfp, M=M)


def mesh_precompute(node_coords, element_vertices, order, face_bc=None) -> dict:
    """Host precompute of the device path (geometry, neighbours, node permutation)."""
    lib = load_library()
    X = np.ascontiguousarray(node_coords, dtype=np.float64)
    V = np.ascontiguousarray(element_vertices, dtype=np.int64)
    K, Np = X.shape[0], X.shape[1]
    dim = V.shape[1] - 1
    desc = MeshDesc(dim, order, K, Np, _ptr(X, C.c_double), _ptr(V, C.c_int64), None)
    if face_bc is not None:
        B = np.ascontiguousarray(face_bc, dtype=np.int8)
        desc.face_bc = _ptr(B, C.c_int8)
    Nfp = order + 1 if dim == 2 else num_nodes(2, order)
    nperm = 2 if dim == 2 else 6
    Ji = np.zeros((K, 3, 3))
    nb = np.zeros((K, dim + 1), dtype=np.int32)
    cd = np.zeros((K, dim + 1), dtype=np.uint8)
    fp = np.zeros((nperm, Nfp), dtype=np.int32)
    tc = np.zeros((K, Np), dtype=np.int32)
    st = lib.dg_mesh_precompute(C.byref(desc), _ptr(Ji, C.c_double), _ptr(nb, C.c_int32), _ptr(cd, C.c_uint8),
                                _ptr(fp, C.c_int32), _ptr(tc, C.c_int32))
    if st != DG_OK:
        raise DGError(st, "dg_mesh_precompute failed")
    return dict(Jinv=Ji, nbr=nb, code=cd, face_perm=fp, to_canon=tc)


# ------------------------------------------------------------ learned closure network ----

def closure_sizes(nfp: int, dim: int = 2) -> tuple[int, int]:
    """(n_in, n_out) of the closure network for dim-dimensional faces with nfp nodes."""
    return dim * nfp + 2, nfp


def closure_num_params(nfp: int, dim: int = 2) -> int:
    nin, nout = closure_sizes(nfp, dim)
    return int(load_library().dg_closure_num_params(nin, nout))


def _closure_nout(nparams: int, nin: int) -> int:
    r = nparams - (64 * nin + 64 + 64 * 64 + 64)
    if r <= 0 or r % 65:
        raise ValueError(f"{nparams} parameters do not fit a network with {nin} inputs")
    return r // 65


def to_bf16(x) -> np.ndarray:
    """FP32 -> BF16 bit patterns (uint16), round to nearest even (library routine)."""
    a = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.zeros(a.size, dtype=np.uint16)
    load_library().dg_closure_to_bf16(_ptr(a, C.c_float), a.size, _ptr(out, C.c_uint16))
    return out


def _closure_check(st):
    if st != DG_OK:
        raise DGError(st, load_library().dg_closure_last_error().decode())


def closure_features(traces, normals, v0, rho0, c0, with_targets=True, dim=2):
    """Features X [n][dim nfp + 2], targets Y [n][nfp] (or None) and scales s [n] of face traces
    [n][2][nfp][nf] (interior, neighbour) of dim-dimensional faces with outward normals [n][3]."""
    T = np.ascontiguousarray(traces, dtype=np.float64)
    Nn = np.ascontiguousarray(normals, dtype=np.float64)
    n, _, nfp, nf_ = T.shape
    nin, nout = closure_sizes(nfp, dim)
    X = np.zeros((n, nin), dtype=np.float32)
    Y = np.zeros((n, nout), dtype=np.float32) if with_targets else None
    S = np.zeros(n, dtype=np.float32)
    ph = Physics((C.c_double * 3)(*v0), rho0, c0, DG_FIELDS_LEE5 if nf_ == 5 else DG_FIELDS_REF4)
    st = load_library().dg_closure_features(dim, nfp, nf_, n, _ptr(T, C.c_double), _ptr(Nn, C.c_double), C.byref(ph),
                                            _ptr(X, C.c_float), _ptr(Y, C.c_float) if Y is not None else None,
                                            _ptr(S, C.c_float))
    if st != DG_OK:
        raise DGError(st, "dg_closure_features: invalid arguments")
    return X, Y, S


def closure_forward(params_bf16, X, device=0) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, nin = X.shape
    W = np.ascontiguousarray(params_bf16, dtype=np.uint16)
    nout = _closure_nout(W.size, nin)
    Y = np.zeros((n, nout), dtype=np.float32)
    _closure_check(load_library().dg_closure_forward(nin, nout, _ptr(W, C.c_uint16), _ptr(X, C.c_float), n,
                                                     _ptr(Y, C.c_float), device))
    return Y


def closure_loss_grad(params_bf16, X, Y, device=0):
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    n, nin = X.shape
    nout = Y.shape[1]
    W = np.ascontiguousarray(params_bf16, dtype=np.uint16)
    g = np.zeros(W.size, dtype=np.float32)
    loss = C.c_double()
    _closure_check(load_library().dg_closure_loss_grad(nin, nout, _ptr(W, C.c_uint16), _ptr(X, C.c_float),
                                                       _ptr(Y, C.c_float), n, C.byref(loss), _ptr(g, C.c_float),
                                                       device))
    return loss.value, g


def closure_init(nfp: int, seed: int = 0, out_scale: float = 0.0, dim: int = 2) -> np.ndarray:
    """FP32 master parameters: Glorot-uniform hidden layers, output layer scaled by
    out_scale (0: the network starts as the zero map, i.e. the characteristic condition)."""
    nin, nout = closure_sizes(nfp, dim)
    rng = np.random.default_rng(seed)
    H = 64

    def glorot(o, i):
        lim = np.sqrt(6.0 / (i + o))
        return rng.uniform(-lim, lim, size=(o, i))

    parts = [glorot(H, nin), np.zeros(H), glorot(H, H), np.zeros(H), out_scale * glorot(nout, H), np.zeros(nout)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def closure_train(master, X, Y, epochs=10, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=1,
                  device=0):
    """Adam on the GPU (the library's forward/backward kernels). Returns (master FP32,
    params BF16, per-epoch mean loss)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    n, nin = X.shape
    nout = Y.shape[1]
    m = np.ascontiguousarray(master, dtype=np.float32).copy()
    W = np.zeros(m.size, dtype=np.uint16)
    hist = np.zeros(max(1, epochs))
    prm = TrainParams(epochs, batch, lr, beta1, beta2, eps, seed)
    _closure_check(load_library().dg_closure_train(nin, nout, _ptr(m, C.c_float), _ptr(W, C.c_uint16),
                                                   _ptr(X, C.c_float), _ptr(Y, C.c_float), n, C.byref(prm),
                                                   _ptr(hist, C.c_double), device))
    return m, W, hist[:epochs]


@dataclass
class RunResult:
    steps: int
    times: np.ndarray
    residuals: np.ndarray
    observers: np.ndarray
    device_ms: float
    saves: list = field(default_factory=list)


class Solver:
    """One solver context on one GPU (dg_create ... dg_destroy)."""

    def __init__(self, node_coords, element_vertices, order, face_bc=None, v0=(0.0, 0.0, 0.0), rho0=1.0,
                 c0=1.0, fields=DG_FIELDS_REF4, precision=DG_FP64, use_graph=True, steps_per_graph=16,
                 device=0):
        self._lib = load_library()
        X = np.ascontiguousarray(node_coords, dtype=np.float64)
        V = np.ascontiguousarray(element_vertices, dtype=np.int64)
        K, Np = X.shape[0], X.shape[1]
        dim = V.shape[1] - 1
        self.dim, self.order, self.K, self.Np = dim, order, K, Np
        self.nf = 5 if fields == DG_FIELDS_LEE5 else 4
        self._keep = [X, V]
        desc = MeshDesc(dim, order, K, Np, _ptr(X, C.c_double), _ptr(V, C.c_int64), None)
        if face_bc is not None:
            B = np.ascontiguousarray(face_bc, dtype=np.int8)
            self._keep.append(B)
            desc.face_bc = _ptr(B, C.c_int8)
        ph = Physics((C.c_double * 3)(*v0), rho0, c0, fields)
        op = Options(precision, 1 if use_graph else 0, steps_per_graph, device)
        h = C.c_void_p()
        st = self._lib.dg_create(C.byref(desc), C.byref(ph), C.byref(op), C.byref(h))
        if st != DG_OK:
            raise DGError(st, self._lib.dg_last_error(None).decode())
        self._h = h
        self._callbacks = []
        self.num_observers = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != DG_OK:
            raise DGError(st, self._lib.dg_last_error(self._h).decode())

    @property
    def state_shape(self):
        return (self.nf, self.K * self.Np)

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.state_shape)
        self._check(self._lib.dg_set_state(self._h, _ptr(u, C.c_double)))

    def get_state(self) -> np.ndarray:
        u = np.zeros(self.state_shape)
        self._check(self._lib.dg_get_state(self._h, _ptr(u, C.c_double)))
        return u

    def add_gaussian(self, x, y, z, size, amp):
        self._check(self._lib.dg_add_gaussian(self._h, x, y, z, size, amp))

    def add_monopole(self, x, y, z, size, amp, freq, phase, duration) -> int:
        n = C.c_int64()
        self._check(self._lib.dg_add_source_monopole(self._h, x, y, z, size, amp, freq, phase, duration,
                                                     C.byref(n)))
        return n.value

    def add_table_source(self, x, y, z, size, fn) -> int:
        """fn(t) -> value or None (None: no overwrite at t)."""
        def cb(_user, t, out):
            v = fn(t)
            if v is None:
                return 0
            out[0] = float(v)
            return 1
        cfn = VALUE_FN(cb)
        self._callbacks.append(cfn)
        n = C.c_int64()
        self._check(self._lib.dg_add_source_table(self._h, x, y, z, size, cfn, None, C.byref(n)))
        return n.value

    def add_observer(self, x, y, z, size) -> int:
        n = C.c_int64()
        self._check(self._lib.dg_add_observer(self._h, x, y, z, size, C.byref(n)))
        self.num_observers += 1
        return n.value

    def set_closure(self, params_bf16):
        """Network of the DG_BC_LEARNED faces (BF16 bit patterns, canonical order)."""
        W = np.ascontiguousarray(params_bf16, dtype=np.uint16)
        self._check(self._lib.dg_set_closure(self._h, _ptr(W, C.c_uint16), W.size))

    def learned_faces(self):
        """(element ids, local faces) of the learned faces in ghost order."""
        n = C.c_int64()
        self._check(self._lib.dg_learned_faces(self._h, None, None, 0, C.byref(n)))
        e = np.zeros(n.value, dtype=np.int64)
        f = np.zeros(n.value, dtype=np.int32)
        self._check(self._lib.dg_learned_faces(self._h, _ptr(e, C.c_int64), _ptr(f, C.c_int32), n.value,
                                               C.byref(n)))
        return e, f

    def face_traces(self, elem, face):
        """Interior/neighbour traces [n][2][Nfp][nf] and outward normals [n][3] of the
        current state on the faces (elem[i], face[i])."""
        e = np.ascontiguousarray(elem, dtype=np.int64)
        f = np.ascontiguousarray(face, dtype=np.int32)
        nfp = self.info().nodes_per_face
        out = np.zeros((e.size, 2, nfp, self.nf))
        nrm = np.zeros((e.size, 3))
        self._check(self._lib.dg_face_traces(self._h, e.size, _ptr(e, C.c_int64), _ptr(f, C.c_int32),
                                             _ptr(out, C.c_double), _ptr(nrm, C.c_double)))
        return out, nrm

    def info(self) -> Info:
        inf = Info()
        self._check(self._lib.dg_get_info(self._h, C.byref(inf)))
        return inf

    def rhs(self, u, dt) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.state_shape)
        k = np.zeros(self.state_shape)
        self._check(self._lib.dg_rhs(self._h, _ptr(u, C.c_double), _ptr(k, C.c_double), dt))
        return k

    def run(self, t_start, t_end, dt, t_rate, integrator=DG_RK4, residuals=True, max_steps=-1,
            num_observers=0, on_save=None) -> RunResult:
        prm = RunParams(t_start, t_end, dt, t_rate, integrator, 1 if residuals else 0, max_steps)
        times, res, obs, saves = [], [], [], []

        def step_cb(_user, recp):
            r = recp.contents
            times.append(r.t)
            res.append([r.residuals[i] for i in range(5)])
            obs.append([r.observers[i] for i in range(4 * r.num_observers)])
            return 0

        def save_cb(_user, step, t, state):
            saves.append((step, t))
            if on_save is not None:
                arr = np.ctypeslib.as_array(state, shape=(self.nf * self.K * self.Np,)).copy()
                on_save(step, t, arr.reshape(self.state_shape))
            return 0

        done = C.c_int64()
        st = self._lib.dg_run(self._h, C.byref(prm), STEP_FN(step_cb), SAVE_FN(save_cb), None, C.byref(done))
        result = RunResult(done.value, np.array(times), np.array(res).reshape(-1, 5),
                           np.array(obs).reshape(len(obs), -1, 4) if obs else np.zeros((0, 0, 4)),
                           self._lib.dg_last_run_device_ms(self._h), saves)
        if st != DG_OK:
            err = DGError(st, self._lib.dg_last_error(self._h).decode())
            err.result = result
            raise err
        return result

    def run_steps(self, nsteps, dt, t_start=0.0, integrator=DG_RK4, residuals=True, capacity=None):
        """Run exactly nsteps steps (no saves after step 0) collecting records into arrays."""
        cap = capacity or nsteps
        prm = RunParams(t_start, t_start + (nsteps + 0.5) * dt, dt, 1e300, integrator, 1 if residuals else 0,
                        nsteps)
        times = np.zeros(cap)
        res = np.zeros((cap, 5))
        obs = np.zeros((cap, max(1, self.num_observers), 4))
        done = C.c_int64()
        self._check(self._lib.dg_run_collect(self._h, C.byref(prm), _ptr(times, C.c_double), _ptr(res, C.c_double),
                                             _ptr(obs, C.c_double), cap, C.byref(done)))
        n = done.value
        return RunResult(n, times[:n], res[:n], obs[:n, :self.num_observers], self._lib.dg_last_run_device_ms(self._h))

    def profile_steps(self, dt, steps):
        st = (C.c_double * 4)()
        src = C.c_double()
        post = C.c_double()
        self._check(self._lib.dg_profile_steps(self._h, dt, steps, st, C.byref(src), C.byref(post)))
        return list(st), src.value, post.value
