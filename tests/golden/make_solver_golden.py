"""Writes solver_recommended.json from the reference's own cache of solver
results (hyperparameter_solver/cache/recommended_values.json): per key
"{mode}_size{N}_W{W}_T{S_trans}_S{S_slot}" the recommended [alpha, b, cost].
Run in the build container (the reference tree is not on the GPU hosts):

    python tests/golden/make_solver_golden.py /root/reference
"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "hyperparameter_solver", "cache", "recommended_values.json")
with open(src) as f:
    entries = json.load(f)["entries"]
out = {k: v for k, v in sorted(entries.items()) if not k.startswith("_")}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "solver_recommended.json")
with open(dst, "w") as f:
    json.dump({"source": "hyperparameter_solver/cache/recommended_values.json", "entries": out}, f, indent=0)
print(f"{len(out)} entries -> {dst}")
