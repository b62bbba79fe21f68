"""Write tests/golden/golden_evolver.json: tnkit's stateful TDVP (`tnkit.tdvp_evolver.TDVPEvolver`)
on the '1site', '2site' and '12site' schedules, recorded by running tnkit itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_evolver.py

Each run starts from tnkit's ground state of the transverse-field Ising chain N=10 at g0 (two-site
DMRG, chi=4, so that the bonds start below the cap), quenches to g1 and takes steps of dt with the
adaptive Krylov exponential (fast_expm=None) or a fixed schedule (fast_expm=m). Recorded per step
(numbers as float.hex): <Z_i> of every site, the bond dimensions, and for '12site' the number of
two-site bonds of the step (tnkit's profile key n_two). Runs split across calls (run(k) twice)
check that the evolver continues where it stopped.
"""
from __future__ import annotations

import json
import os
import sys

for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.environ.get("TNKIT_REF", "/root/reference"))

from tnkit.api import dmrg  # noqa: E402
from tnkit.models import tfim_mpo  # noqa: E402
from tnkit.observables import local_expectations  # noqa: E402
from tnkit.tdvp_evolver import TDVPEvolver  # noqa: E402

Z = np.diag([1.0, -1.0])


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).reshape(-1)]


def run_case(name, method, chi, n_steps, split, g0=1.5, g1=0.8, dt=0.05, fast_expm=None,
             two_site_every=0, svd_tol=0.0, degenerate_tol=0.0, N=10):
    _, gs, _ = dmrg(tfim_mpo(N, g0), 4, n_sweeps=4, tol=0.0, rtol=0.0)
    gs = [np.asarray(t, dtype=np.complex128) for t in gs]
    steps = []

    def cb(d, rec):
        steps.append({"n_two": d.get("n_two"), "max_chi": d.get("max_chi")})

    ev = TDVPEvolver(gs, tfim_mpo(N, g1), dt, chi=chi, backend="numpy", fast_expm=fast_expm,
                     svd_tol=svd_tol, degenerate_tol=degenerate_tol, step_callback=cb,
                     measure=lambda m: local_expectations(m, Z), meas_every=1)
    recs = []
    for part in split:
        recs += ev.run(part, method=method, two_site_every=two_site_every)
    assert sum(split) == n_steps
    bonds = [int(t.shape[2]) for t in ev.mps[:-1]]
    return {"name": name, "method": method, "chi": chi, "n_steps": n_steps, "split": split,
            "g0": g0, "g1": g1, "dt": dt, "N": N, "fast_expm": fast_expm,
            "two_site_every": two_site_every, "svd_tol": svd_tol,
            "degenerate_tol": degenerate_tol,
            "z": [hexs(z) for _, z in recs], "t": [float(t).hex() for t, _ in recs],
            "n_two": [s["n_two"] for s in steps], "max_chi": [s["max_chi"] for s in steps],
            "final_bonds": bonds}


def main():
    runs = [
        run_case("2site", "2site", 8, 6, [6]),
        run_case("1site", "1site", 8, 6, [6]),
        run_case("12site", "12site", 8, 8, [8]),
        run_case("12site-split", "12site", 8, 8, [3, 5]),
        run_case("12site-every", "12site", 8, 8, [8], two_site_every=3),
        run_case("12site-fixed", "12site", 8, 6, [6], fast_expm=10),
        run_case("12site-svdtol", "12site", 16, 6, [6], svd_tol=1e-10),
    ]
    with open(os.path.join(HERE, "golden_evolver.json"), "w") as f:
        json.dump({"runs": runs}, f, indent=1)
    for r in runs:
        print(r["name"], r["n_two"], r["max_chi"], r["final_bonds"])


if __name__ == "__main__":
    main()
