"""The dgalerkin command-line driver on the reference's own 3D and second-order meshes, from
.conf files, against the oracle run on the same mesh (read by the independent Python reader).

  sphere:  doc/3d/sphere.msh, 20,211 linear tetrahedra in a ball of radius 5, physical surface
           'Absorbing'; RK4, dt = 2e-5 to t = 2e-3 (100 steps), v0 = (20, 0, 0), Gaussian
           initial pressure at (1, 0, 0), one monopole source, two observers; snapshot at step 50.
  square2: doc/2d/square2.msh, 3,706 second-order (6-node) triangles, physical curve 'Abs'
           mapped to reflecting walls; forward Euler (Euler1), dt = 5e-6 to t = 5e-4
           (100 steps), v0 = (20, -10, 0), Gaussian initial pressure, a dipole source (two
           monopoles in antiphase), one observer; snapshot at step 50.
  square2_tables: the same mesh, absorbing, RK4 over 100 steps with a quadrupole, a formula
           source overlapping it (later in key order, so it wins while active; its duration
           ends at step 60) and a file source (data/data.csv, linear interpolation of the
           float-rounded table, include/configParser.h:38-58), two observers.

The CLI prints residuals and observer values with 6 significant digits, so those are held to
rtol 1e-5; VTU cell averages (17 digits) to 1e-10 of the field's maximum.
"""
import csv
import os
import shutil
import subprocess

import numpy as np
import pytest

import dgfem_amd as dg
from msh_py import read_msh41

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "ref_cases")
EXE = os.path.join(HERE, "..", "dgfem-acoustic_amd", "dgalerkin")
RHO0, C0 = 1.225, 340.0

CONF = """[Config]
meshFileName = {mesh}
timeStart=0
timeEnd={t_end}
timeStep={dt}
timeRate={rate}
elementType=Lagrange
timeIntMethod={method}
{bc}
numThreads=1
v0_x = {v0[0]}
v0_y = {v0[1]}
v0_z = {v0[2]}
rho0 = {rho0}
c0 = {c0}
{extra}
"""

CASE_DEFS = {
    "sphere": dict(mesh="doc/3d/sphere.msh", t_end=2e-3, dt=2e-5, rate=1e-3, method="Runge-Kutta",
                   bc="Absorbing = Absorbing", v0=(20.0, 0.0, 0.0), integrator=0, bc_type=0,
                   extra="initialCondtition1 = gaussian, 1,0,0, 0.5, 1\n"
                         "source1 = monopole, -1,1,0, 0.4, 0.5, 300, 0, 1\n"
                         "observer1 = 2,0,0,0.3\nobserver2 = 0,-2,1,0.8\n",
                   ic=(1.0, 0.0, 0.0, 0.5, 1.0), observers=[(2.0, 0.0, 0.0, 0.3), (0.0, -2.0, 1.0, 0.8)]),
    "square2": dict(mesh="doc/2d/square2.msh", t_end=5e-4, dt=5e-6, rate=2.5e-4, method="Euler1",
                    bc="Abs = Reflecting", v0=(20.0, -10.0, 0.0), integrator=1, bc_type=1,
                    extra="initialCondtition1 = gaussian, 1,1,0, 2, 1\n"
                          "source1 = dipole, -2,0,0, 0.6, 0.3, 2000, 0, 1\n"
                          "observer1 = 3,0,0,0.5\n",
                    ic=(1.0, 1.0, 0.0, 2.0, 1.0), observers=[(3.0, 0.0, 0.0, 0.5)]),
    "square2_tables": dict(mesh="doc/2d/square2.msh", files=["data/data.csv"], t_end=5e-4, dt=5e-6, rate=2.5e-4,
                           method="Runge-Kutta", bc="Abs = Absorbing", v0=(20.0, -10.0, 0.0), integrator=0, bc_type=0,
                           extra="initialCondtition1 = gaussian, 1,1,0, 2, 1\n"
                                 "source1 = quadrupole, -2,0,0, 0.8, 0.4, 2500, 0, 1\n"
                                 "source2 = formula, \"0.2*sin(2*pi*3000*t)+0.05\", -2,0,0, 0.5, 3.0e-4\n"
                                 "source3 = file,\"data/data.csv\", 1.5,-1,0, 0.5\n"
                                 "observer1 = 3,0,0,0.5\nobserver2 = -1.5,0.5,0,0.6\n",
                           ic=(1.0, 1.0, 0.0, 2.0, 1.0), observers=[(3.0, 0.0, 0.0, 0.5), (-1.5, 0.5, 0.0, 0.6)]),
}


def _interpolate(rows, t):
    """Sources::interpolate_value (include/configParser.h:38-58): linear between the bracketing
    rows found by bisection, 0 outside [t_first, t_last]."""
    if len(rows) < 2 or not rows[0, 0] <= t <= rows[-1, 0]:
        return 0.0
    lo = min(int(np.searchsorted(rows[:, 0], t, side="right")) - 1, len(rows) - 2)
    hi = lo + 1
    return rows[lo, 1] + ((rows[hi, 1] - rows[lo, 1]) / (rows[hi, 0] - rows[lo, 0])) * (t - rows[lo, 0])


def _oracle_sources(c, summary):
    """Oracle source list (kind 0 monopoles; kind 1 for formula and file sources, whose values
    come from the table function) and the table function."""
    src, tables = [], {}
    for i, x in enumerate(summary["sources"]):
        if x["kind"] == 0:
            src.append(x)
            continue
        src.append(dict(kind=1, x=x["x"], y=x["y"], z=x["z"], size=x["size"]))
        if x["kind"] == 1:
            tables[i] = (lambda t, f=x["formula"], d=x["duration"]: dg.eval_formula(f, t) if t < d else None)
        else:
            (path,) = c["files"]
            rows = dg.read_source_table(os.path.join(CASES, path))
            assert len(rows) == x["rows"]
            tables[i] = (lambda t, r=rows: _interpolate(r, t))
    return src, (lambda i, t: tables[i](t))


def _read_residuals(path):
    with open(path) as f:
        rows = list(csv.reader(f, delimiter=";"))
    assert rows[0] == ["time", "res_p", "res_rho", "res_vx", "res_vy", "res_vz", "elapsed_time"]
    return np.array([[float(x) for x in r[:6]] for r in rows[1:]])


@pytest.fixture(scope="module", params=sorted(CASE_DEFS))
def case(request, tmp_path_factory):
    name = request.param
    c = CASE_DEFS[name]
    if not os.path.exists(EXE):
        pytest.fail("dgalerkin is not built (make -C dgfem-acoustic_amd)")
    d = tmp_path_factory.mktemp(name)
    for rel in [c["mesh"]] + c.get("files", []):
        os.makedirs(d / os.path.dirname(rel), exist_ok=True)
        shutil.copyfile(os.path.join(CASES, rel), d / rel)
    (d / "case.conf").write_text(CONF.format(rho0=RHO0, c0=C0, **c))
    cwd = os.getcwd()
    try:
        os.chdir(d)
        summary = dg.case_summary("case.conf")
    finally:
        os.chdir(cwd)
    print(f"case cli {name} conf fp64", flush=True)
    r = subprocess.run([EXE, "case.conf"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr

    import oracle_py as orc
    m = read_msh41(os.path.join(CASES, c["mesh"]))
    X = m["coords"]
    (tag,) = m["groups"]
    o = orc.Oracle(X, m["verts"], summary["order"], [m["groups"][tag]], [c["bc_type"]], v0=c["v0"], rho0=RHO0,
                   c0=C0)
    xs = X.reshape(-1, 3)
    x0, y0, z0, size, amp = c["ic"]
    u0 = np.zeros((4, len(xs)))
    u0[0] = amp * np.exp(-((xs[:, 0] - x0) ** 2 + (xs[:, 1] - y0) ** 2 + (xs[:, 2] - z0) ** 2) / size)
    src, table_fn = _oracle_sources(c, summary)
    run = dict(integrator=c["integrator"], sources=src, table_fn=table_fn, observers=c["observers"])
    full = o.run(u0, 0.0, c["t_end"], c["dt"], c["rate"], **run)
    assert np.isfinite(full["observers"]).all()
    half = o.run(u0, 0.0, c["t_end"], c["dt"], c["rate"], max_steps=int(full["saves"][1]), **run)
    return {**c, "name": name, "dir": d, "stdout": r.stdout, "summary": summary, "msh": m, "full": full, "half": half}


def test_case_shape(case):
    s, m = case["summary"], case["msh"]
    assert s["dim"] == m["dim"] and s["elements"] == len(m["tags"])
    if case["name"] == "square2":
        assert s["order"] == 2 and s["bc_counts"] == {"absorbing": 0, "reflecting": 160}
        assert [x["pole"] for x in s["sources"]] == [1, 1]
        assert [x["phase"] for x in s["sources"]] == [0.0, np.pi]
    elif case["name"] == "square2_tables":
        assert s["order"] == 2 and s["bc_counts"] == {"absorbing": 160, "reflecting": 0}
        assert [x["kind"] for x in s["sources"]] == [0, 0, 0, 0, 1, 2]
        assert [x["pole"] for x in s["sources"]][:4] == [2, 2, 2, 2]
    else:
        assert s["order"] == 1 and s["bc_counts"] == {"absorbing": 3198, "reflecting": 0}
    assert case["full"]["steps"] == 100


def test_cli_residuals_match_oracle(case):
    res = _read_residuals(case["dir"] / "residuals.csv")
    full = case["full"]
    assert len(res) == full["steps"]
    np.testing.assert_allclose(res[:, 0], full["times"], rtol=1e-5, atol=1e-15)
    for k in range(5 if case["msh"]["dim"] == 3 else 4):
        np.testing.assert_allclose(res[:, 1 + k], full["residuals"][:, k], rtol=1e-5, atol=0, err_msg=f"column {k}")
    assert np.abs(full["residuals"][:, 0]).max() > 0
    if case["msh"]["dim"] == 2:
        assert np.all(res[:, 5] == 0.0)


def test_cli_observers_match_oracle(case):
    full = case["full"]
    for j in range(len(case["observers"])):
        rows = open(case["dir"] / "results" / f"observers{j + 1}.txt").read().splitlines()
        # forward Euler's header names pressure before density, but its rows are written
        # density first like RK4's (src/solver.cpp:150 and :290, :409 and :572)
        head = "time;pressure;density;" if case["integrator"] == 1 else "time;density;pressure;"
        assert rows[0] == head + "velocity_x;velocity_y;velocity_z"
        vals = np.array([[float(x) for x in r.split(";")] for r in rows[1:]])
        assert vals.shape == (full["steps"], 6)
        ob = full["observers"][:, j, :]
        for col, k in ((2, 0), (3, 1), (4, 2), (5, 3))[:case["msh"]["dim"] + 1]:
            sc = np.abs(ob[:, k]).max()
            np.testing.assert_allclose(vals[:, col], ob[:, k], rtol=1e-5, atol=1e-9 * sc, err_msg=f"obs {j} col {col}")
        np.testing.assert_allclose(vals[:, 1], ob[:, 0] / C0 ** 2, rtol=1e-5, atol=1e-9 * np.abs(ob[:, 0]).max() / C0 ** 2)


def test_cli_snapshot_matches_oracle(case):
    half = case["half"]
    step = half["steps"]
    K, Np = case["msh"]["coords"].shape[:2]
    assert f"Step number : {step}" in case["stdout"]
    vtu = open(case["dir"] / "results" / f"result{step}.vtu").read()
    assert f'NumberOfCells="{K}"' in vtu

    def cell_array(name):
        body = vtu.split(f'Name="{name}"')[1].split(">", 1)[1].split("</DataArray>")[0]
        return np.array(body.split(), dtype=float)
    u = half["u"]
    # cell value = mean over the element's Np nodes (src/Mesh.cpp:1194-1210)
    p_ref = u[0].reshape(K, Np).mean(axis=1)
    v_ref = np.stack([u[1 + c].reshape(K, Np).mean(axis=1) for c in range(3)], axis=1)
    sc = np.abs(p_ref).max()
    assert sc > 0
    assert np.abs(cell_array("Pressure [Pa]") - p_ref).max() < 1e-10 * sc
    vs = np.abs(v_ref).max()
    assert np.abs(cell_array("Velocity [m/s]").reshape(K, 3) - v_ref).max() < 1e-10 * vs


def test_cli_mixed_precision_close_to_fp64(case):
    """DGFEM_PRECISION=mixed selects the mixed-precision path in the command-line driver: its
    observer series stay within 1e-5 relative L2 of the FP64 run's (both printed with 6
    significant digits)."""
    d = case["dir"]
    mixed = d / "mixed"
    if not mixed.exists():
        shutil.copytree(d, mixed, ignore=shutil.ignore_patterns("results", "residuals.csv"))
    print(f"case cli {case['name']} conf mixed", flush=True)
    env = dict(os.environ, DGFEM_PRECISION="mixed")
    r = subprocess.run([EXE, "case.conf"], cwd=mixed, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    for j in range(len(case["observers"])):
        def series(path):
            rows = open(path).read().splitlines()[1:]
            return np.array([[float(x) for x in row.split(";")] for row in rows])
        a = series(d / "results" / f"observers{j + 1}.txt")
        b = series(mixed / "results" / f"observers{j + 1}.txt")
        assert a.shape == b.shape
        assert np.array_equal(a[:, 0], b[:, 0])
        col = 2  # pressure (rows are time; density; pressure; ... for both integrators)
        assert np.linalg.norm(b[:, col] - a[:, col]) <= 1e-5 * np.linalg.norm(a[:, col])
