"""Front-end (host) tests: formula sources, file-source tables, configuration files and
mesh ingest, against the reference's own evaluator output and data files (fixtures under
tests/golden/, written by tests/golden/make_fixtures.py)."""
import json
import os
import shutil
import struct
import subprocess
import wave

import numpy as np
import pytest

import dgfem_amd as dg

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = os.path.join(GOLD, "ref_cases")
REF_EVAL = os.path.join(HERE, "..", "oracle", "_ref", "eqedit_ref")


def test_formula_matches_reference_evaluator_corpus():
    corpus = json.load(open(os.path.join(GOLD, "eqedit_corpus.json")))
    assert len(corpus) > 200
    for item in corpus:
        ref = float(item["value"])
        mine = dg.eval_formula(item["expr"], item["t"])
        assert (mine == ref) or (np.isnan(mine) and np.isnan(ref)), (item, mine)


@pytest.mark.skipif(not os.path.exists(REF_EVAL), reason="reference evaluator not built (oracle/build_ref.sh)")
def test_formula_matches_reference_evaluator_random_expressions():
    rng = np.random.default_rng(0)
    atoms = ["t", "pi", "e", "2", "0.5", "3e-2", "sin(t)", "exp(-t)", "abs(t-1)", "(t+1)", "ln(2+t)"]
    ops = ["+", "-", "*", "/", "^"]
    exprs = []
    for _ in range(300):
        n = rng.integers(1, 6)
        s = ("-" if rng.random() < 0.2 else "") + rng.choice(atoms)
        for _ in range(n):
            s += rng.choice(ops) + rng.choice(atoms)
        if rng.random() < 0.3:
            s = "(" + s + ")*" + rng.choice(atoms)
        exprs.append(s)
    ts = [0.0, 0.37, 1.9]
    pairs = [(e, t) for e in exprs for t in ts]
    out = subprocess.run([REF_EVAL], input="".join(f"{e}\t{t!r}\n" for e, t in pairs), capture_output=True,
                         text=True, check=True).stdout.split()
    for (e, t), r in zip(pairs, out):
        ref, mine = float(r), dg.eval_formula(e, t)
        assert (mine == ref) or (np.isnan(mine) and np.isnan(ref)), (e, t, mine, ref)


def test_csv_source_table_rounds_through_float():
    path = os.path.join(CASES, "data", "data.csv")
    rows = dg.read_source_table(path)
    lines = [l.strip() for l in open(path)][1:]
    lines = [l for l in lines if l and not l.startswith("#")]
    expect = np.array([[np.float32(float(x)) for x in l.split(";")] for l in lines], dtype=np.float64)
    assert rows.shape == expect.shape
    assert np.array_equal(rows, expect)
    assert not np.array_equal(rows[:, 1], np.array([float(l.split(";")[1]) for l in lines]))


def test_wav_source_table_scaling_and_times():
    path = os.path.join(CASES, "data", "data1.wav")
    rows = dg.read_source_table(path)
    with wave.open(path) as w:
        nch, width, rate, nfr = w.getnchannels(), w.getsampwidth(), w.getframerate(), w.getnframes()
        raw = w.readframes(nfr)
    assert width == 2
    ints = np.array(struct.unpack("<%dh" % (len(raw) // 2), raw))
    vals = (ints.astype(np.float32) / np.float32(32767)).astype(np.float64)
    assert rows.shape[0] == nfr * nch
    assert np.array_equal(rows[:, 1], vals)
    t = 0.0
    step = 1.0 / rate
    for i in range(min(50, len(rows))):
        assert rows[i, 0] == t
        t += step


def _stage_case(tmp_path, *rels):
    for rel in rels:
        dst = tmp_path / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(os.path.join(CASES, rel), dst)


def test_hilltop_bc_counts_match_first_node_rule():
    """The C++ reader + BC rule against an independent Python reader + the same rule
    (src/Mesh.cpp:579-605): abs (tag 1) absorbing, then ref (tag 2) reflecting."""
    from msh_py import first_node_bc, read_msh41_tri
    X, tris, groups = read_msh41_tri(os.path.join(CASES, "cases", "2d_hilltop", "2d_hilltop_fs.msh"))
    bc = first_node_bc(tris, groups, {1: 0, 2: 1})
    cwd = os.getcwd()
    try:
        os.chdir(CASES)
        js = dg.case_summary("cases/2d_hilltop/2d_hilltop.json")
    finally:
        os.chdir(cwd)
    assert js["elements"] == len(tris)
    assert js["bc_counts"]["reflecting"] == sum(1 for v in bc.values() if v == 1)
    assert js["bc_counts"]["absorbing"] == sum(1 for v in bc.values() if v == 0)


def test_hilltop_json_and_conf_cases(tmp_path, monkeypatch):
    _stage_case(tmp_path, "cases/2d_hilltop/2d_hilltop.json", "cases/2d_hilltop/2d_hilltop.conf",
                "cases/2d_hilltop/2d_hilltop_fs.msh")
    monkeypatch.chdir(tmp_path)
    js = dg.case_summary("cases/2d_hilltop/2d_hilltop.json")
    assert js["dim"] == 2 and js["order"] == 1 and js["elements"] == 7199
    assert js["time"] == [0.0, 1.0, 1e-3, 0.1]
    assert js["integrator"] == "Runge-Kutta" and js["num_threads"] == 20
    assert js["mean_flow"] == [0.0, 0.0, 0.0, 1.225, 340.0]
    assert js["initial_conditions"] == [[0, 0.0, 200.0, 0.0, 2000.0, 1.0]]
    assert js["observers"] == [[500.0, 2.0, 0.0, 10.0]]
    assert js["sources"] == []
    assert sorted(t for _, t in js["physical_bcs"]) == ["Absorbing", "Reflecting"]
    nb = js["bc_counts"]["absorbing"] + js["bc_counts"]["reflecting"]
    assert js["bc_counts"]["reflecting"] > 0 and js["bc_counts"]["absorbing"] > 0
    cf = dg.case_summary("cases/2d_hilltop/2d_hilltop.conf")
    assert cf["bc_counts"] == js["bc_counts"] and cf["physical_bcs"] == js["physical_bcs"]
    assert cf["bc_counts"]["absorbing"] + cf["bc_counts"]["reflecting"] == nb
    assert cf["time"] == [0.0, 1.5, 1e-4, 0.001] and cf["num_threads"] == 64
    src = cf["sources"]
    assert len(src) == 1 and src[0]["kind"] == 0 and src[0]["pole"] == 0
    assert (src[0]["x"], src[0]["y"], src[0]["size"], src[0]["amp"], src[0]["freq"]) == (0.0, 200.0, 10.0, 2.0, 5.0)


CONF = """[Config]
# comment
meshFileName = doc/2d/square_reflection.msh
timeStart=0
timeEnd=0.01
timeStep=1e-4
timeRate=0.005
elementType=Lagrange
timeIntMethod=Euler1
Absorbing = Absorbing
Reflecting = ReflectingWall
numThreads=1
v0_x = 10
v0_y = -5
v0_z = 0
rho0 = 1.2
c0 = 343
source2 = dipole, 1,2,0, 0.4, 2,10,0.5,5
source10 = quadrupole, -1,0,0, 0.6, 3,20,0,1
source3 = formula, "0.1 * sin(2 * pi * 50 * t)", 0.5,0.5,0.0, 0.1, 0.2
source4 = file,"data/data.csv", 0.0,0.0,0.0, 0.1
observer1 = 1,1,0,0.5
initialCondtition1 = gaussian, 0,0,0,1,2
"""


def test_conf_grammar_sources_and_bcs(tmp_path, monkeypatch):
    _stage_case(tmp_path, "doc/2d/square_reflection.msh", "data/data.csv")
    (tmp_path / "case.conf").write_text(CONF)
    monkeypatch.chdir(tmp_path)
    c = dg.case_summary("case.conf")
    assert c["integrator"] == "Euler1" and c["num_threads"] == 0
    assert c["mean_flow"] == [10.0, -5.0, 0.0, 1.2, 343.0]
    s = c["sources"]
    # keys visited in lexicographic order: source10 < source2 < source3 < source4
    assert [x["pole"] for x in s] == [2, 2, 2, 2, 1, 1, -1, -1]
    q = s[:4]
    assert [(x["x"], x["y"]) for x in q] == [(-1.6, 0.0), (-0.4, 0.0), (-1.0, -0.6), (-1.0, 0.6)]
    assert all(x["size"] == 0.3 for x in q)
    assert [x["phase"] for x in q] == [0.0, 0.0, np.pi, np.pi]
    d = s[4:6]
    assert [(x["x"], x["size"], x["phase"]) for x in d] == [(0.6, 0.2, 0.5), (1.4, 0.2, 0.5 + np.pi)]
    assert s[6]["kind"] == 1 and s[6]["formula"] == "0.1*sin(2*pi*50*t)" and s[6]["duration"] == 0.2
    assert s[7]["kind"] == 2 and s[7]["rows"] == 51
    assert c["initial_conditions"] == [[0, 0.0, 0.0, 0.0, 1.0, 2.0]]
    # 'Reflecting = ReflectingWall' matches by prefix; both physical curves are typed
    assert sorted(t for _, t in c["physical_bcs"]) == ["Absorbing", "Reflecting"]
    assert c["bc_counts"]["reflecting"] > 0 and c["bc_counts"]["absorbing"] > 0


def test_missing_files_report_reference_messages(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with pytest.raises(dg.DGError) as e:
        dg.case_summary("nope.conf")
    assert "Project file 'nope.conf' not found or corrupted" in str(e.value)
    (tmp_path / "m.conf").write_text(CONF.replace("doc/2d/square_reflection.msh", "missing.msh"))
    (tmp_path / "data").mkdir()
    shutil.copyfile(os.path.join(CASES, "data", "data.csv"), tmp_path / "data" / "data.csv")
    with pytest.raises(dg.DGError) as e:
        dg.case_summary("m.conf")
    assert "Mesh file 'missing.msh' not found or corrupted" in str(e.value)


MESH_CONF = """[Config]
meshFileName = {mesh}
timeStart=0
timeEnd=0.01
timeStep=1e-4
timeRate=0.005
elementType=Lagrange
timeIntMethod=Runge-Kutta
{bcs}
numThreads=1
v0_x = 0
v0_y = 0
v0_z = 0
rho0 = 1.225
c0 = 340
"""


def _mesh_summary(tmp_path, monkeypatch, rel, bcs):
    _stage_case(tmp_path, rel)
    (tmp_path / "m.conf").write_text(MESH_CONF.format(mesh=rel, bcs="\n".join(bcs)))
    monkeypatch.chdir(tmp_path)
    return dg.case_summary("m.conf")


@pytest.mark.parametrize("rel,tag_types,order,npe", [
    ("doc/2d/square2.msh", {9: 0}, 2, 6),     # second-order triangles, physical 'Abs'
    ("doc/3d/sphere.msh", {1: 0}, 1, 4),      # linear tetrahedra, physical surface 'Absorbing'
    ("doc/2d/square.msh", {9: 0}, 2, 6),      # second-order triangles, physical 'Absorbing'
    ("doc/2d/disk.msh", {2: 1}, 1, 3),        # linear triangles, physical 'Reflecting'
])
def test_reference_meshes_against_python_reader(tmp_path, monkeypatch, rel, tag_types, order, npe):
    """Element count, order, node count per element and boundary faces of the reference's own
    triangle meshes (order 1 and 2) and tetrahedral mesh, against the independent Python reader
    and the first-node rule (src/Mesh.cpp:579-605)."""
    from msh_py import first_node_bc_any, read_msh41
    m = read_msh41(os.path.join(CASES, rel))
    c = _mesh_summary(tmp_path, monkeypatch, rel, ["Abs = Absorbing", "Absorbing = Absorbing",
                                                   "Reflecting = Reflecting"])
    assert c["dim"] == m["dim"] and c["order"] == order and c["nodes_per_element"] == npe
    assert c["elements"] == len(m["tags"]) and m["tags"].shape[1] == npe
    bc = first_node_bc_any(m["verts"], m["dim"], m["groups"], tag_types)
    assert c["bc_counts"]["absorbing"] == sum(1 for v in bc.values() if v == 0)
    assert c["bc_counts"]["reflecting"] == sum(1 for v in bc.values() if v == 1)
    assert [n for _, n in c["physical_bcs"]] == [("Absorbing", "Reflecting")[tag_types[t]] for t in sorted(tag_types)]
    # a closed boundary: every boundary face is typed and the count is exactly the unpaired faces
    assert len(bc) == sum(c["bc_counts"].values()) > 0


def test_reference_example_configs(tmp_path, monkeypatch):
    """The reference's own example configuration in both grammars (doc/config/Square.conf and
    Square.json, on doc/2d/square2.msh)."""
    _stage_case(tmp_path, "doc/config/Square.conf", "doc/config/Square.json", "doc/2d/square2.msh")
    monkeypatch.chdir(tmp_path)
    c = dg.case_summary("doc/config/Square.conf")
    j = dg.case_summary("doc/config/Square.json")
    for x, t_end, threads, duration in ((c, 0.1, 12, 0.1), (j, 0.05, 16, 0.05)):
        assert x["dim"] == 2 and x["order"] == 2 and x["elements"] == 3706
        assert x["time"] == [0.0, t_end, 5e-05, 0.001]
        assert x["integrator"] == "Runge-Kutta" and x["num_threads"] == threads
        assert x["mean_flow"] == [30.0, 0.0, 0.0, 1.225, 100.0]
        assert x["bc_counts"] == {"absorbing": 160, "reflecting": 0}
        assert x["observers"] == [[2.11792, 0.00340081, 0.0, 0.1]]
        assert x["initial_conditions"] == []
        (s,) = x["sources"]
        assert s["kind"] == 1 and s["formula"] == "0.1*sin(2*pi*50*t)" and s["duration"] == duration
        assert (s["x"], s["y"], s["z"], s["size"]) == (0.0, 0.0, 0.0, 0.1)


def test_reference_msh40_mesh(tmp_path, monkeypatch):
    """The reference's room mesh is written in MSH 4.0 (header '4 0 8'): the reader takes it
    and its triangle count matches the file's element blocks."""
    rel = "doc/2d/room.msh"
    L = open(os.path.join(CASES, rel)).read().split("\n")
    assert L[1].split()[:2] == ["4", "0"]
    i = L.index("$Elements")
    j, ntri = i + 2, 0
    for _ in range(int(L[i + 1].split()[0])):
        a = list(map(int, L[j].split()))   # 4.0 block header: entity tag, dim, type, count
        if a[1] == 2 and a[2] == 2:
            ntri += a[3]
        j += 1 + a[3]
    c = _mesh_summary(tmp_path, monkeypatch, rel, ["Absorbing = Absorbing"])
    assert c["dim"] == 2 and c["order"] == 1 and c["elements"] == ntri == 5199
    assert c["bc_counts"] == {"absorbing": 451, "reflecting": 0}
