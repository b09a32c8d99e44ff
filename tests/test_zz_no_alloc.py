"""dg_face_traces_noalloc: the face-trace entry point without device allocation (staging in a
fixed per-device array), against dg_face_traces on the same faces, bit for bit, on two calls in
a row (the second smaller), in FP64 and mixed precision; a request larger than the array is
rejected on the host."""
import ctypes as C

import numpy as np
import pytest

import dgfem_amd as dg
from cases import gaussian_state, make_case

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 8 << 20


def _align(n):
    return (n + 255) // 256 * 256


def _staging_bytes(n, nfp, nf):
    return _align(8 * n) + _align(4 * n) + _align(16 * n * nfp * nf) + _align(24 * n)


def _noalloc(s, elem, face):
    lib = dg.load_library()
    e = np.ascontiguousarray(elem, dtype=np.int64)
    f = np.ascontiguousarray(face, dtype=np.int32)
    nfp = s.info().nodes_per_face
    out = np.full((e.size, 2, nfp, s.nf), np.nan)
    nrm = np.full((e.size, 3), np.nan)
    st = lib.dg_face_traces_noalloc(s._h, e.size, e.ctypes.data_as(C.POINTER(C.c_int64)),
                                    f.ctypes.data_as(C.POINTER(C.c_int32)), out.ctypes.data_as(C.POINTER(C.c_double)),
                                    nrm.ctypes.data_as(C.POINTER(C.c_double)))
    return st, out, nrm


@pytest.mark.parametrize("precision", [dg.DG_FP64, dg.DG_MIXED])
def test_noalloc_traces_equal_the_allocating_entry_point(precision):
    print(f"case face traces noalloc precision={precision}", flush=True)
    groups = [(0x10, dg.DG_BC_REFLECTING), (0x2F, dg.DG_BC_ABSORBING)]
    m, bc, _ = make_case(3, 6, 3, groups, jitter=0.15, seed=31)
    s = dg.Solver(m.node_coords, m.element_vertices, 3, bc, v0=(30.0, 0.0, 0.0), rho0=1.225, c0=340.0,
                  precision=precision)
    s.set_state(gaussian_state(m, 0.1, -0.2, 0.3, 0.05) + 0.01)
    rng = np.random.default_rng(5)
    for n in (5000, 1234):  # the second call smaller than the first
        elem = rng.integers(0, m.K, n)
        face = rng.integers(0, 4, n).astype(np.int32)
        ref, nref = s.face_traces(elem, face)
        st, out, nrm = _noalloc(s, elem, face)
        assert st == dg.DG_OK
        np.testing.assert_array_equal(out, ref)
        np.testing.assert_array_equal(nrm, nref)
        assert np.abs(out).max() > 0
    # beyond the staging array: rejected before any device work
    nfp, nf = s.info().nodes_per_face, s.nf
    nmax = SCRATCH_BYTES // (16 * nfp * nf + 36)
    while _staging_bytes(nmax + 1, nfp, nf) <= SCRATCH_BYTES:
        nmax += 1
    while _staging_bytes(nmax, nfp, nf) > SCRATCH_BYTES:
        nmax -= 1
    big = nmax + 1
    st, _, _ = _noalloc(s, np.zeros(big, np.int64), np.zeros(big, np.int32))
    assert st == dg.DG_ERR_INVALID
    assert "staging" in s._lib.dg_last_error(s._h).decode()
    st, out, _ = _noalloc(s, np.zeros(nmax, np.int64), np.zeros(nmax, np.int32))
    assert st == dg.DG_OK and np.isfinite(out).all()
    s.close()
