"""The oracle against golden outputs of the reference's own Fortran
(tests/golden/ref_*.json, written by tests/golden/make_golden.py from the
reference compiled unchanged).  Runs without the reference build: inputs
are regenerated from their seeds and every output must match the
reference's bytes (SHA-256)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from golden_cases import CASES, world_for  # noqa: E402
from oracle import restate as O  # noqa: E402

GOLD = os.path.join(HERE, "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def gold_fn():
    with open(os.path.join(GOLD, "ref_functions.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("cid", sorted(CASES))
def test_function_golden(gold_fn, cid):
    fname, names, touched, setup = CASES[cid]
    w = world_for(cid)
    setup(w, None)
    kw = {"v_rot": 0.3, "r_rot": 2.0} if cid == "bndry_inner" else {}
    getattr(O, fname)(*w.args(names), **kw)
    bad = []
    for k in touched:
        g = gold_fn[f"{cid}/{k}"]
        if sha(w.a[k]) != g["sha256"]:
            bad.append((k, float(w.a[k].astype(np.float64).sum()), g["sum"]))
    assert not bad, bad


def test_tiny_steps_golden():
    from g3m3_amd.workload import make_workload
    from oracle.boxstep import OracleRun, FIELDS
    with open(os.path.join(GOLD, "ref_tiny_steps.json")) as f:
        gold = json.load(f)
    r = OracleRun(make_workload("tiny"), "restate")
    r.init_dt()
    for s in (1, 2):
        r.full_step()
        assert [float(x) for x in r.t_step] == gold[f"step{s}/t_step"]
        bad = [k for k in FIELDS if sha(r.a[k]) != gold[f"step{s}/{k}"]["sha256"]]
        assert not bad, (s, bad)
