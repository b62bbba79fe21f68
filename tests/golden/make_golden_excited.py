"""Write tests/golden/excited_golden.json from tnkit's recorded excited-state runs.

tnkit ships its dense `dmrg_excited` / `dmrg_tower` results as a data file
(tests/data/excited_tower_golden.json, written by tnkit's tests/test_excited_symmetric.py:90-117):
transverse-field Ising chain N=8, g=1.2; the previous state is `dmrg(W, chi=8, n_sweeps=4)`;
`dmrg_excited(W, 8, prev_states=[gs], n_sweeps=6, seed=3, **kw)` per run; the towers are
`dmrg_tower(W, 16, 5, n_sweeps=8, **kw)`. Numbers are float.hex strings.

This script copies the numbers (not the state hashes, which are bytes of one machine's
arithmetic) and records the calls that made them. It reads the data file from the read-only
reference mount and runs nothing from it.

    python tests/golden/make_golden_excited.py
"""
from __future__ import annotations

import json
import os

SRC = "/root/reference/tests/data/excited_tower_golden.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "excited_golden.json")

CALLS = {
    "excited-penalty": {"kw": {}, "dtype": "float64"},
    "excited-complex": {"kw": {}, "dtype": "complex128"},
    "excited-project": {"kw": {"deflate": "project"}, "dtype": "float64"},
    "excited-lanczos": {"kw": {"eig": "lanczos"}, "dtype": "float64"},
    "excited-init": {"kw": {"init_gs": True, "penalty": 2.0}, "dtype": "float64"},
    "tower-penalty": {"kw": {}, "dtype": "float64"},
    "tower-project": {"kw": {"deflate": "project"}, "dtype": "float64"},
}


def main() -> None:
    with open(SRC) as f:
        gold = json.load(f)
    runs = {}
    for name, call in CALLS.items():
        rec = dict(gold["runs"][name])
        rec.pop("state", None)
        rec.pop("states", None)
        runs[name] = {"call": call, "record": rec}
    out = {"model": {"kind": "tfim", "N": 8, "g": 1.2},
           "ground_state": {"chi": 8, "n_sweeps": 4},
           "excited": {"chi": 8, "n_sweeps": 6, "seed": 3},
           "tower": {"chi": 16, "nlevels": 5, "n_sweeps": 8},
           "recorded_on": gold["machine"], "runs": runs}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
