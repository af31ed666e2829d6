"""Writes tests/golden/ref_*.npz: outputs of the reference's own Fortran
subroutines (oracle/_ref, compiled unchanged from the reference sources) on
seeded inputs, so the oracle stays pinned where that build is absent.

    python tests/golden/make_golden.py      (needs oracle/_ref built)

Inputs are not stored: they are regenerated from the seeds by
tests/common.World and g3m3_amd.workload.make_workload.  Each output array
is recorded by the SHA-256 of its float32 bytes (bit-exact identity) and its
float64 sum (for diagnosing a mismatch).
"""


def digest(a):
    a = np.ascontiguousarray(a)
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "sum": float(a.astype(np.float64).sum()), "shape": list(a.shape),
            "dtype": str(a.dtype)}
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.dirname(os.path.dirname(HERE))]

from oracle import ref as R  # noqa: E402
from oracle.boxstep import OracleRun, FIELDS  # noqa: E402
from golden_cases import CASES, world_for  # noqa: E402
from g3m3_amd.workload import make_workload  # noqa: E402


def main():
    lib = R.RefLib(omp=False)
    out = {}
    for cid, (fname, names, touched, setup) in CASES.items():
        w = world_for(cid)
        setup(w, lib)
        lib.call(fname, *w.args(names))
        for k in touched:
            out[f"{cid}/{k}"] = digest(w.a[k])
    with open(os.path.join(HERE, "ref_functions.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    # two full steps of the two-box workload, whole state
    wl = make_workload("tiny")
    r = OracleRun(wl, "ref")
    r.init_dt()
    st = {}
    for s in (1, 2):
        r.full_step()
        for k in FIELDS:
            st[f"step{s}/{k}"] = digest(r.a[k])
        st[f"step{s}/t_step"] = [float(x) for x in r.t_step]
    with open(os.path.join(HERE, "ref_tiny_steps.json"), "w") as f:
        json.dump(st, f, indent=0, sort_keys=True)
    print("wrote", len(out), "function outputs and", len(st), "step arrays")


if __name__ == "__main__":
    main()
