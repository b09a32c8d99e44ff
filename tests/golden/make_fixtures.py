"""Writes small fixtures derived from data files held in the reference repository.

Run in the build container (where the reference checkout exists); the outputs are
committed and the tests read only them:
  results_pvd.json   - save schedule of the 2d_hilltop run (results.pvd:3-12) and the loop
                       parameters it was produced with (cases/2d_hilltop/2d_hilltop.json:19-25)
  ref_cases/         - copies of data files of the reference (example case configs, meshes,
                       source tables, the 2d_hilltop residual history)
  eqedit_corpus.json - values of the reference formula evaluator (src/eqEdit.cpp, built by
                       oracle/build_ref.sh into oracle/_ref/eqedit_ref) on an expression corpus
"""
import json
import os
import re
import shutil
import subprocess
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def results_pvd():
    text = open(os.path.join(REF, "results.pvd")).read()
    sets = re.findall(r'<DataSet timestep="([^"]+)" part="0" file="results/result(\d+)\.vtu"/>', text)
    case = json.load(open(os.path.join(REF, "cases", "2d_hilltop", "2d_hilltop.json")))
    t = case["solver"]["time"]
    out = {"t_start": t["start"], "t_end": t["end"], "dt": t["step"], "t_rate": t["rate"],
           "saves": [{"step": int(s), "timestep": ts} for ts, s in sets],
           "num_residual_rows": sum(1 for _ in open(os.path.join(REF, "residuals.csv"))) - 1}
    json.dump(out, open(os.path.join(HERE, "results_pvd.json"), "w"), indent=1)


DATA_FILES = [
    "cases/2d_hilltop/2d_hilltop.json",
    "cases/2d_hilltop/2d_hilltop.conf",
    "cases/2d_hilltop/2d_hilltop_fs.msh",
    "doc/2d/square_reflection.msh",
    "doc/2d/square2.msh",
    "doc/2d/room.msh",
    "doc/3d/sphere.msh",
    "doc/2d/square.msh",
    "doc/2d/disk.msh",
    "doc/config/Square.conf",
    "doc/config/Square.json",
    "data/data.csv",
    "data/data1.wav",
    "residuals.csv",
    "results.pvd",
]

CORPUS = [
    "0.1*sin(2*pi*50*t)", "2^3^2", "-2^2", "2*-3", "1+2", "exp(-t)*t", "(1+2", "1+2)", "1)+2", "-(3-5)*2",
    "2(3)", "abs(-3.5)+int(2.7)+frac(2.7)+rond(2.5)+rond(-2.5)+rond(2.4)", "log(100)+ln(e)", "sh(1)+ch(1)+th(1)",
    "asin(0.5)+acos(0.5)+atan(1)", "ash(-2)+ash(2)+ach(2)+ath(0.5)", "t^2-3*t/4+1", "2e-3*t", "1e3", "pi*e", "-t",
    "--t", "t*t*t", "(t)", "((t+1)*(t-1))/2", "sin(t", "cos()", "foo(1)", "x+1", "0.5*(1+tan(t))", "2^-1",
    "3/0", "-0", "exp(710)", "sqrt(2)", "1.5.2", "7%2", "t2", "2t", "(-1)^0.5", "-(-1)", "1-(-1)", "1--1",
    "4^0.5^2", "-t^2", "-(t^2)", "2*(3+4)^2/7", "sin(cos(tan(t)))", "int(-t)", "frac(-t)", "rond(t)",
]
TIMES = [0.0, 0.013, 0.5, 1.7, -0.25]


def ref_cases():
    out = os.path.join(HERE, "ref_cases")
    for rel in DATA_FILES:
        dst = os.path.join(out, rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(os.path.join(REF, rel), dst)


def eqedit_corpus():
    exe = os.path.join(HERE, "..", "..", "oracle", "_ref", "eqedit_ref")
    pairs = [(e, t) for e in CORPUS for t in TIMES]
    text = "".join(f"{e}\t{t!r}\n" for e, t in pairs)
    res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split()
    json.dump([{"expr": e, "t": t, "value": v} for (e, t), v in zip(pairs, res)],
              open(os.path.join(HERE, "eqedit_corpus.json"), "w"), indent=0)


if __name__ == "__main__":
    results_pvd()
    ref_cases()
    eqedit_corpus()
