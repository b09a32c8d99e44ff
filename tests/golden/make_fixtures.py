"""Writes small fixtures derived from data files held in the reference repository.

Run in the build container (where the reference checkout exists); the outputs are
committed and the tests read only them:
  results_pvd.json  - save schedule of the 2d_hilltop run (results.pvd:3-12) and the loop
                      parameters it was produced with (cases/2d_hilltop/2d_hilltop.json:19-25)
"""
import json
import os
import re
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


if __name__ == "__main__":
    results_pvd()
