"""Writes two fixtures from the reference's own solver outputs (data files,
read as text; no reference code is run):

* solver_recommended.json, from hyperparameter_solver/cache/recommended_values.json:
  per key "{mode}_size{N}_W{W}_T{S_trans}_S{S_slot}" the recommended
  [alpha, b, cost];
* solver_grid.json, from hyperparameter_solver/res/sc18.csv, sc19.csv, sc20.csv:
  the full 19 x 9 (alpha, b) grid of costs of each dataset (table type N;
  the H rows carry the same costs), as [alpha, b, cost] rows.  The costs are
  printed to 6 decimals; the files do not record the table size or the
  hardware parameters.

Run in the build container (the reference tree is not on the GPU hosts):

    python tests/golden/make_solver_golden.py /root/reference
"""
import csv
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))

src = os.path.join(ref, "hyperparameter_solver", "cache", "recommended_values.json")
with open(src) as f:
    entries = json.load(f)["entries"]
out = {k: v for k, v in sorted(entries.items()) if not k.startswith("_")}
dst = os.path.join(here, "solver_recommended.json")
with open(dst, "w") as f:
    json.dump({"source": "hyperparameter_solver/cache/recommended_values.json", "entries": out}, f, indent=0)
print(f"{len(out)} entries -> {dst}")

grids = {}
for sc in ("sc18", "sc19", "sc20"):
    with open(os.path.join(ref, "hyperparameter_solver", "res", sc + ".csv")) as f:
        rows = list(csv.DictReader(f))
    by = {}
    for r in rows:
        by.setdefault(r["table_type"], []).append([float(r["alpha"]), int(r["b"]), float(r["cost"])])
    assert sorted(x[:2] + [x[2]] for x in by["N"]) == sorted(x[:2] + [x[2]] for x in by["H"])
    grids[sc] = by["N"]
dst = os.path.join(here, "solver_grid.json")
with open(dst, "w") as f:
    json.dump({"source": "hyperparameter_solver/res/sc{18,19,20}.csv (table_type N)", "grids": grids}, f,
              separators=(",", ":"))
print(f"{sum(len(g) for g in grids.values())} grid rows -> {dst}")
