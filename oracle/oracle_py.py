"""ctypes binding of the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by the test suite, the smoke check and the CPU-baseline leg of the benchmark;
never by the product library. See dg_oracle.h for what is restated from the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("order", C.c_int), ("num_elements", C.c_int64), ("nodes_per_element", C.c_int),
                ("node_coords", C.POINTER(C.c_double)), ("elem_verts", C.POINTER(C.c_int64)),
                ("num_groups", C.c_int), ("group_type", C.POINTER(C.c_int8)),
                ("group_offsets", C.POINTER(C.c_int64)), ("group_nodes", C.POINTER(C.c_int64))]


class OPhysics(C.Structure):
    _fields_ = [("v0", C.c_double * 3), ("rho0", C.c_double), ("c0", C.c_double)]


class OSource(C.Structure):
    _fields_ = [("kind", C.c_int), ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("size", C.c_double),
                ("amp", C.c_double), ("freq", C.c_double), ("phase", C.c_double), ("duration", C.c_double)]


class OObserver(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("size", C.c_double)]


VALUE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double))


class ORunParams(C.Structure):
    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("dt", C.c_double), ("t_rate", C.c_double),
                ("integrator", C.c_int), ("num_sources", C.c_int), ("sources", C.POINTER(OSource)),
                ("value_fn", VALUE_FN), ("value_user", C.c_void_p), ("num_observers", C.c_int),
                ("observers", C.POINTER(OObserver)), ("max_steps", C.c_int64), ("num_threads", C.c_int)]


class ORunOutput(C.Structure):
    _fields_ = [("steps", C.c_int64), ("times", C.POINTER(C.c_double)), ("residuals", C.POINTER(C.c_double)),
                ("observers", C.POINTER(C.c_double)), ("save_steps", C.POINTER(C.c_int64)),
                ("num_saves", C.c_int64)]


_lib = None


def load(build_if_missing: bool = True) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) and build_if_missing:
        subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.oracle_create.restype = C.c_void_p
    lib.oracle_create.argtypes = [P(OMesh), P(OPhysics)]
    lib.oracle_destroy.argtypes = [C.c_void_p]
    lib.oracle_error.restype = C.c_char_p
    lib.oracle_rhs.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_double, C.c_int]
    lib.oracle_run.argtypes = [C.c_void_p, P(C.c_double), P(ORunParams), P(ORunOutput)]
    lib.oracle_num_volume_points.argtypes = [C.c_void_p]
    lib.oracle_num_face_points.argtypes = [C.c_void_p]
    lib.oracle_num_faces.argtypes = [C.c_void_p]
    lib.oracle_num_faces.restype = C.c_int64
    lib.oracle_face_bc.argtypes = [C.c_void_p, C.c_int64, C.c_int]
    lib.oracle_bc_flux.argtypes = [P(OPhysics), C.c_int] + [P(C.c_double)] * 5
    _lib = lib
    return lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def group_node_sets(element_vertices: np.ndarray, face_side: np.ndarray, groups) -> list[np.ndarray]:
    """Vertex sets of physical groups defined by generator side masks (Gmsh local face tables)."""
    dim = element_vertices.shape[1] - 1
    fv = [[0, 1], [1, 2], [2, 0]] if dim == 2 else [[0, 2, 1], [0, 1, 3], [0, 3, 2], [3, 1, 2]]
    out = []
    for mask, _typ in groups:
        verts = []
        for f in range(dim + 1):
            side = face_side[:, f].astype(np.int64)
            sel = (side >= 0) & ((np.uint32(mask) >> np.clip(side, 0, 31).astype(np.uint32)) & 1).astype(bool)
            verts.append(element_vertices[sel][:, fv[f]].ravel())
        out.append(np.unique(np.concatenate(verts)) if verts else np.zeros(0, np.int64))
    return out


def bc_flux(bc, n, t, s, q, v0=(0, 0, 0), rho0=1.0, c0=1.0):
    """Boundary flux of the oracle at one face point (bc 0 absorbing, 1 reflecting)."""
    lib = load()
    ph = OPhysics((C.c_double * 3)(*v0), rho0, c0)
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (n, t, s, q)]
    out = np.zeros(4)
    lib.oracle_bc_flux(C.byref(ph), bc, *[_p(a, C.c_double) for a in arr], _p(out, C.c_double))
    return out


class Oracle:
    def __init__(self, node_coords, element_vertices, order, group_nodes=(), group_types=(), v0=(0, 0, 0),
                 rho0=1.0, c0=1.0):
        self._lib = load()
        X = np.ascontiguousarray(node_coords, dtype=np.float64)
        V = np.ascontiguousarray(element_vertices, dtype=np.int64)
        self.K, self.Np = X.shape[0], X.shape[1]
        self.dim = V.shape[1] - 1
        offs = np.zeros(len(group_nodes) + 1, dtype=np.int64)
        for g, s in enumerate(group_nodes):
            offs[g + 1] = offs[g] + len(s)
        nodes = np.concatenate([np.sort(np.asarray(s, np.int64)) for s in group_nodes]) if group_nodes else \
            np.zeros(1, np.int64)
        types = np.array(list(group_types) or [0], dtype=np.int8)
        self._keep = [X, V, offs, nodes, types]
        m = OMesh(self.dim, order, self.K, self.Np, _p(X, C.c_double), _p(V, C.c_int64), len(group_nodes),
                  _p(types, C.c_int8), _p(offs, C.c_int64), _p(nodes, C.c_int64))
        ph = OPhysics((C.c_double * 3)(*v0), rho0, c0)
        self._h = self._lib.oracle_create(C.byref(m), C.byref(ph))
        if not self._h:
            raise RuntimeError("oracle_create: " + self._lib.oracle_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_volume_points(self):
        return self._lib.oracle_num_volume_points(self._h)

    @property
    def num_faces(self):
        return self._lib.oracle_num_faces(self._h)

    def face_bc(self):
        """Boundary types [K][dim+1] per local face: 0 absorbing, 1 reflecting, -1 interior."""
        K, nfc = self.K, self.dim + 1
        return np.array([[self._lib.oracle_face_bc(self._h, e, f) for f in range(nfc)] for e in range(K)])

    def rhs(self, u, dt, threads=0):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(4, self.K * self.Np)
        k = np.zeros_like(u)
        self._lib.oracle_rhs(self._h, _p(u, C.c_double), _p(k, C.c_double), dt, threads)
        return k

    def run(self, u, t_start, t_end, dt, t_rate, integrator=0, sources=(), table_fn=None, observers=(),
            max_steps=None, threads=0):
        """sources: dicts with kind 0 (monopole: x,y,z,size,amp,freq,phase,duration) or kind 1
        (x,y,z,size; values from table_fn(source_index, t) -> value or None)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(4, self.K * self.Np).copy()
        if max_steps is None:
            max_steps = int(np.floor((t_end - t_start) / dt)) + 2
        src = (OSource * max(1, len(sources)))()
        for i, s in enumerate(sources):
            src[i] = OSource(s.get("kind", 0), s["x"], s["y"], s["z"], s["size"], s.get("amp", 0), s.get("freq", 0),
                             s.get("phase", 0), s.get("duration", 0))
        obs = (OObserver * max(1, len(observers)))()
        for i, o in enumerate(observers):
            obs[i] = OObserver(*o)

        def vf(_u, i, t, out):
            if table_fn is None:
                return 0
            v = table_fn(i, t)
            if v is None:
                return 0
            out[0] = float(v)
            return 1

        cvf = VALUE_FN(vf)
        prm = ORunParams(t_start, t_end, dt, t_rate, integrator, len(sources), src, cvf, None, len(observers), obs,
                         max_steps, threads)
        times = np.zeros(max_steps)
        res = np.zeros((max_steps, 5))
        ob = np.zeros((max_steps, max(1, len(observers)), 4))
        saves = np.zeros(max_steps, dtype=np.int64)
        out = ORunOutput(0, _p(times, C.c_double), _p(res, C.c_double), _p(ob, C.c_double), _p(saves, C.c_int64), 0)
        self._lib.oracle_run(self._h, _p(u, C.c_double), C.byref(prm), C.byref(out))
        n = out.steps
        return dict(u=u, steps=n, times=times[:n], residuals=res[:n], observers=ob[:n, :len(observers)],
                    saves=saves[:out.num_saves])
