"""Independent Python reader of the Gmsh 4.1 ASCII meshes the tests use (triangles of order 1
and 2, linear tetrahedra, boundary physical groups), so that the C++ reader and the BC rule are
checked against a second implementation. Test infrastructure only."""
from __future__ import annotations

import numpy as np

TRI_EDGES = ((0, 1), (1, 2), (2, 0))  # Gmsh local edges of a triangle; first node = first entry


def read_msh41_tri(path):
    L = open(path).read().split("\n")
    i = L.index("$Entities")
    npt, ncu, _, _ = map(int, L[i + 1].split())
    curve_phys = {}
    for k in range(ncu):
        a = L[i + 2 + npt + k].split()
        nph = int(a[7])
        curve_phys[int(a[0])] = [int(x) for x in a[8:8 + nph]]
    i = L.index("$Nodes")
    nb = int(L[i + 1].split()[0])
    j, nodes = i + 2, {}
    for _ in range(nb):
        n = int(L[j].split()[3])
        j += 1
        tags = [int(L[j + k]) for k in range(n)]
        j += n
        for k in range(n):
            nodes[tags[k]] = [float(x) for x in L[j + k].split()[:3]]
        j += n
    i = L.index("$Elements")
    nb = int(L[i + 1].split()[0])
    j, tris, groups = i + 2, [], {}
    for _ in range(nb):
        d, t, ty, n = map(int, L[j].split())
        j += 1
        for k in range(n):
            ids = [int(x) for x in L[j + k].split()[1:]]
            if d == 2 and ty == 2:
                tris.append(ids)
            elif d == 1:
                for ph in curve_phys.get(t, []):
                    groups.setdefault(ph, set()).update(ids)
        j += n
    tris = np.array(tris, dtype=np.int64)
    X = np.array([[nodes[v] for v in t] for t in tris])
    return X, tris, {ph: np.array(sorted(s), dtype=np.int64) for ph, s in groups.items()}


def first_node_bc(tris, groups, types):
    """BC type per boundary edge by the first-node rule: default absorbing (0); groups in
    ascending tag order, later groups overwrite. Returns {(element, local edge): type}."""
    count = {}
    for e, t in enumerate(tris):
        for f, (a, b) in enumerate(TRI_EDGES):
            count.setdefault(tuple(sorted((t[a], t[b]))), []).append((e, f))
    bc = {}
    for key, owners in count.items():
        if len(owners) == 1:
            bc[owners[0]] = 0
    for ph in sorted(groups):
        s = set(groups[ph].tolist())
        for (e, f) in bc:
            if tris[e][TRI_EDGES[f][0]] in s:
                bc[(e, f)] = types[ph]
    return bc


ELEMENT_TYPES = {2: (2, 3), 9: (2, 6), 4: (3, 4), 11: (3, 10)}  # gmsh type -> (dim, nodes)
TET_FACES = ((0, 2, 1), (0, 1, 3), (0, 3, 2), (3, 1, 2))  # Gmsh local faces; first node = first entry


def read_msh41(path):
    """Top-dimension elements of a Gmsh 4.1 ASCII mesh: node coordinates [K][Np][3] and node
    tags [K][Np] in file order, the dimension, and the node sets of the boundary physical groups
    ({physical tag: sorted node tags}) together with their names."""
    L = open(path).read().split("\n")
    names = {}
    if "$PhysicalNames" in L:
        i = L.index("$PhysicalNames")
        for k in range(int(L[i + 1])):
            a = L[i + 2 + k].split()
            names[(int(a[0]), int(a[1]))] = a[2].strip('"')
    i = L.index("$Entities")
    counts = list(map(int, L[i + 1].split()))
    ent_phys = {}
    j = i + 2
    for d in range(4):
        for _ in range(counts[d]):
            a = L[j].split()
            j += 1
            at = 4 if d == 0 else 7
            nph = int(a[at])
            ent_phys[(d, int(a[0]))] = [int(x) for x in a[at + 1:at + 1 + nph]]
    i = L.index("$Nodes")
    nb = int(L[i + 1].split()[0])
    j, nodes = i + 2, {}
    for _ in range(nb):
        n = int(L[j].split()[3])
        j += 1
        tags = [int(L[j + k]) for k in range(n)]
        j += n
        for k in range(n):
            nodes[tags[k]] = [float(x) for x in L[j + k].split()[:3]]
        j += n
    i = L.index("$Elements")
    nb = int(L[i + 1].split()[0])
    j, blocks = i + 2, []
    for _ in range(nb):
        d, t, ty, n = map(int, L[j].split())
        j += 1
        blocks.append((d, t, ty, [list(map(int, L[j + k].split()[1:])) for k in range(n)]))
        j += n
    dim = max(ELEMENT_TYPES[ty][0] for d, t, ty, e in blocks if ty in ELEMENT_TYPES)
    conn, groups = [], {}
    for d, t, ty, els in blocks:
        if d == dim and ty in ELEMENT_TYPES:
            conn.extend(els)
        elif d == dim - 1:
            for ph in ent_phys.get((d, t), []):
                groups.setdefault(ph, set()).update(x for e in els for x in e)
    conn = np.array(conn, dtype=np.int64)
    X = np.array([[nodes[v] for v in e] for e in conn])
    return dict(dim=dim, coords=X, tags=conn, verts=np.ascontiguousarray(conn[:, :dim + 1]),
                groups={ph: np.array(sorted(v), dtype=np.int64) for ph, v in groups.items()},
                names={ph: names.get((dim - 1, ph), "") for ph in groups})


def first_node_bc_any(verts, dim, groups, types):
    """The first-node rule for triangles or tetrahedra given their vertex tags [K][dim+1]:
    boundary faces default to absorbing (0); groups in ascending tag order overwrite the faces
    whose first local node lies in the group (src/Mesh.cpp:579-605).
    Returns {(element, local face): type}."""
    faces = TRI_EDGES if dim == 2 else TET_FACES
    count = {}
    for e, v in enumerate(verts.tolist()):
        for f, fv in enumerate(faces):
            count.setdefault(tuple(sorted(v[i] for i in fv)), []).append((e, f))
    bc = {owners[0]: 0 for owners in count.values() if len(owners) == 1}
    for ph in sorted(groups):
        s = set(groups[ph].tolist())
        for (e, f) in bc:
            if verts[e][faces[f][0]] in s:
                bc[(e, f)] = types[ph]
    return bc
