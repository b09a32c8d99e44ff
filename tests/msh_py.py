"""Independent Python reader of the Gmsh 4.1 ASCII meshes the tests use (linear triangles,
physical curves), so that the C++ reader and the BC rule are checked against a second
implementation. Test infrastructure only."""
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
