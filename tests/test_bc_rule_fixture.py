"""Hand-labelled boundary types (the reference's first-node rule, src/Mesh.cpp:579-605).

The reference types a boundary face by its FIRST node (the first vertex of the face in Gmsh's
local face table: triangle edges {0,1},{1,2},{2,0}; tetrahedron faces {0,2,1},{0,1,3},{0,3,2},
{3,1,2}) and the physical groups in ascending tag order: every group containing that node sets
the type, later groups overwrite earlier ones, default absorbing. Membership is by node, not by
face, so a face of no group can become reflecting and a face of a reflecting group absorbing.

The meshes under tests/golden/bc_rule were written by hand, and so were the expected types:

square2.msh: nodes 1 (0,0), 2 (1,0), 3 (1,1), 4 (0,1); triangles A = (1,2,3), B = (1,3,4);
  group 1 "wall" (Reflecting) = the left edge (nodes {4,1}), group 2 "open" (Absorbing) = the top
  edge (nodes {3,4}).
  A edge 0 = (1,2), the bottom, in no group: first node 1 is in "wall" only -> reflecting.
  A edge 1 = (2,3), the right side, in no group: node 2 is in no group -> absorbing.
  B edge 1 = (3,4), the top ("open"): node 3 is in "open" only -> absorbing.
  B edge 2 = (4,1), the left side ("wall"): node 4 is in "wall", then in "open" -> absorbing.
tet1.msh: one tetrahedron (1,2,3,4); group 1 "wall" (Reflecting) = face (1,3,2), group 2
  "open" (Absorbing) = face (4,2,3).
  faces 0 (1,3,2), 1 (1,2,4), 2 (1,4,3): first node 1 is in "wall" only -> reflecting (faces 1
  and 2 belong to no group); face 3 (4,2,3): node 4 is in "open" only -> absorbing.
"""
import json
import os

import numpy as np
import pytest

import dgfem_amd as dg
import oracle_py as orc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bc_rule")

# (element in file order, local face, type): 1 reflecting, 0 absorbing
EXPECTED = {
    "square2": [(0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 2, 0)],
    "tet1": [(0, 0, 1), (0, 1, 1), (0, 2, 1), (0, 3, 0)],
}
# the same meshes as arrays: vertex tags per element, coordinates, group vertex sets (tag order)
MESHES = {
    "square2": dict(verts=[[1, 2, 3], [1, 3, 4]], xy={1: (0, 0, 0), 2: (1, 0, 0), 3: (1, 1, 0), 4: (0, 1, 0)},
                    groups=[([4, 1], 1), ([3, 4], 0)]),
    "tet1": dict(verts=[[1, 2, 3, 4]], xy={1: (0, 0, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (0, 0, 1)},
                 groups=[([1, 2, 3], 1), ([2, 3, 4], 0)]),
}


@pytest.mark.parametrize("name", ["square2", "tet1"])
def test_case_reader_types_faces_by_first_node(name, monkeypatch):
    monkeypatch.chdir(HERE)  # the configuration names its mesh relative to the working directory
    js = dg.case_summary(name + ".json")
    got = sorted(tuple(x) for x in js["boundary_faces"])
    assert got == EXPECTED[name]
    assert js["bc_counts"]["reflecting"] == sum(t for *_, t in EXPECTED[name])


@pytest.mark.parametrize("name", ["square2", "tet1"])
def test_oracle_types_faces_by_first_node(name):
    m = MESHES[name]
    V = np.array(m["verts"], dtype=np.int64)
    X = np.array([[m["xy"][v] for v in row] for row in m["verts"]], dtype=np.float64)
    o = orc.Oracle(X, V, 1, [np.array(g) for g, _ in m["groups"]], [t for _, t in m["groups"]])
    bc = o.face_bc()
    got = sorted((e, f, int(bc[e, f])) for e in range(bc.shape[0]) for f in range(bc.shape[1]) if bc[e, f] >= 0)
    assert got == EXPECTED[name]


@pytest.mark.parametrize("name", ["square2", "tet1"])
def test_library_side_rule_types_faces_by_first_node(name):
    """dg_resolve_bc_sides (groups given as side labels) on the same meshes: each boundary face
    is its own side; a group's side mask selects its faces."""
    m = MESHES[name]
    dim = len(m["verts"][0]) - 1
    V = np.array(m["verts"], dtype=np.int64)
    X = np.array([[m["xy"][v] for v in row] for row in m["verts"]], dtype=np.float64)
    # side label per (element, local face): the face's index among the boundary faces
    faces = [(e, f) for e, f, _ in EXPECTED[name]]
    S = -np.ones(V.shape, dtype=np.int8)
    for k, (e, f) in enumerate(faces):
        S[e, f] = k
    fv = [[0, 1], [1, 2], [2, 0]] if dim == 2 else [[0, 2, 1], [0, 1, 3], [0, 3, 2], [3, 1, 2]]
    groups = []
    for gverts, typ in m["groups"]:
        # the group's faces: the boundary faces whose vertices all belong to it
        mask = sum(1 << k for k, (e, f) in enumerate(faces) if set(V[e, fv[f]]) <= set(gverts))
        groups.append((mask, typ))
    mesh = dg.BoxMesh(dim, 1, 0, X, V, S)
    bc = dg.resolve_bc_sides(mesh, groups)
    got = sorted((e, f, int(bc[e, f])) for e, f in faces)
    assert got == EXPECTED[name]
