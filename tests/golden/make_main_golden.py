"""Runs the reference's whole program, unchanged, on the committed input
files and records what it wrote, as golden data for the step sequence.

    make -C oracle main                       # builds oracle/_ref/g3m3_main
    python tests/golden/make_main_golden.py   # writes tests/golden/main_*.json
    python tests/golden/make_main_golden.py --devmath pin_tilt_torus

The program (code/3d_main.f90 with every code/mod_*, sub_* and cft_* file,
built serial by oracle/Makefile) runs in oracle/_ref/main_run/<name>/ with
an unlimited stack.  It reads ``input`` there, asks git for a commit hash
(code/3d_main.f90:659-663), so the directory is made a one-commit git
repository first.  From its output this script keeps, per fluid file
(code/3d_main.f90:3300-3351): the SHA-256 of every record's payload, and for
the 37 field records their float64 sum and sum of squares and the values at
fixed sample cells (lines through the planet in boxes 1 and 2 plus 256
seeded cells); the box-step records of speeds.dat (:2167-2168); and the
wall fluxes of fluxes.dat (:1997-2011).  The fluid files themselves
(661 MB each) are deleted once digested.

--devmath NAME writes main_NAME_devmath.json in the same form from the
oracle instead (tests/mainpin.py:oracle_main_loop: the same set-up, the
reference's compiled subroutines, the restated initialisation block and
torus source), with the initialisation's real power and the torus's
exponentials evaluated in double precision and rounded once, as the device
evaluates them.  The GPU run of NAME is held to it bit for bit; with the C
library's powf / expf the same oracle reproduces the reference program's
own files bit for bit (tests/test_main_pin.py).
"""
from __future__ import annotations

import base64
import hashlib
import json
import os
import re
import resource
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "g3m3_main")
RUNS = os.path.join(ROOT, "oracle", "_ref", "main_run")
INPUTS = ("pin_base", "pin_tilt_torus", "pin_graph")

NX, NY, NZ, NGRD = 121, 121, 61, 5        # the program's compile-time grid
V = NX * NY * NZ
FIELD_RECORDS = ["qrho", "qpx", "qpy", "qpz", "qpresx", "qpresy", "qpresz", "qpresxy",
                 "qpresxz", "qpresyz", "hrho", "hpx", "hpy", "hpz", "hpresx", "hpresy",
                 "hpresz", "hpresxy", "hpresxz", "hpresyz", "orho", "opx", "opy", "opz",
                 "opresx", "opresy", "opresz", "opresxy", "opresxz", "opresyz", "bx", "by",
                 "bz", "epres", "bx0", "by0", "bz0"]
RECORD_NAMES = ["t"] + FIELD_RECORDS + ["lists"]


def sample_index() -> np.ndarray:
    """Flat indices into a (n_grids, nz, ny, nx) array: the x and y lines
    through the box centre (the planet) in boxes 1 and 2, and 256 seeded
    cells anywhere."""
    idx = []
    for b in range(2):
        base = b * V + (NZ // 2) * NX * NY
        idx += [base + (NY // 2) * NX + i for i in range(NX)]
        idx += [base + j * NX + NX // 2 for j in range(NY)]
    rng = np.random.default_rng(20261207)
    idx += list(rng.integers(0, NGRD * V, 256))
    return np.array(idx, np.int64)


def fortran_records(path: str):
    """Payloads of a Fortran unformatted sequential file (4-byte markers)."""
    out = []
    with open(path, "rb") as f:
        while True:
            head = f.read(4)
            if not head:
                break
            n = int(np.frombuffer(head, "<i4")[0])
            data = f.read(n)
            tail = int(np.frombuffer(f.read(4), "<i4")[0])
            if tail != n or len(data) != n:
                raise ValueError(f"{path}: bad record markers")
            out.append(data)
    return out


def digest_fluid(path: str) -> dict:
    recs = fortran_records(path)
    if len(recs) != len(RECORD_NAMES):
        raise ValueError(f"{path}: {len(recs)} records")
    idx = sample_index()
    rows = []
    for name, data in zip(RECORD_NAMES, recs):
        row = {"name": name, "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
        if name == "t":
            row["value"] = float(np.frombuffer(data, "<f8")[0])
        elif name != "lists":
            a = np.frombuffer(data, "<f4")
            a64 = a.astype(np.float64)
            row["sum"] = float(a64.sum())
            row["sumsq"] = float((a64 * a64).sum())
            row["samples"] = base64.b64encode(a[idx].astype("<f4").tobytes()).decode()
        rows.append(row)
    return {"t": rows[0]["value"], "records": rows}


def digest_graph(path: str, every: int = 211) -> dict:
    """A plotting table (code/sub_write_graphing_data.f90:278-760): its
    SHA-256 and line count, the header and column-title lines verbatim, and
    every `every`-th data row (zlib-compressed text)."""
    import zlib
    with open(path, "rb") as f:
        raw = f.read()
    lines = raw.decode().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    head = 6                                     # 5 header lines and the titles
    rows = {str(r): lines[head + r] for r in range(0, len(lines) - head, every)}
    blob = zlib.compress(json.dumps(rows).encode(), 9)
    return {"sha256": hashlib.sha256(raw).hexdigest(), "lines": len(lines),
            "head": lines[:head], "every": every,
            "rows": base64.b64encode(blob).decode()}


_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?"


def parse_speeds(path: str):
    """(box, t, t_old, t_new) of every 'lores t step' record; a record may
    be wrapped over two lines by the list-directed write."""
    text = open(path).read()
    out = []
    for chunk in text.split("lores t step. box/t/old/new:")[1:]:
        v = re.findall(_NUM, chunk)[:4]
        out.append([int(v[0]), float(v[1]), float(v[2]), float(v[3])])
    return out


def parse_fluxes(path: str):
    """[ut, box, q, h, o, 'grid_in'|'grid_out'] of fluxes.dat."""
    out = []
    for line in open(path):
        m = re.match(r"\s*(" + _NUM + r")\s+(\d+)\s+(" + _NUM + r")\s+(" + _NUM + r")\s+("
                     + _NUM + r")\s+(grid_in|grid_out)", line)
        if m:
            out.append([float(m.group(1)), int(m.group(2)), float(m.group(3)),
                        float(m.group(4)), float(m.group(5)), m.group(6)])
    return out


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_one(name: str) -> dict:
    src = os.path.join(HERE, "main_inputs", name + ".input")
    wd = os.path.join(RUNS, name)
    if os.path.isdir(wd):
        shutil.rmtree(wd)
    os.makedirs(wd)
    shutil.copy(src, os.path.join(wd, "input"))
    os.makedirs(os.path.join(wd, "figures", "data"))     # write_graphing_data's directory
    env = dict(os.environ, OMP_NUM_THREADS="1", GIT_AUTHOR_NAME="g", GIT_AUTHOR_EMAIL="g@g",
               GIT_COMMITTER_NAME="g", GIT_COMMITTER_EMAIL="g@g")
    for cmd in (["git", "init", "-q", "."], ["git", "add", "input"],
                ["git", "commit", "-q", "-m", "input"]):
        subprocess.run(cmd, cwd=wd, env=env, check=True)
    print(f"running the reference program on {name} ...", flush=True)
    with open(os.path.join(wd, "stdout.txt"), "w") as so:
        subprocess.run([BIN], cwd=wd, env=env, stdout=so, stderr=subprocess.STDOUT,
                       preexec_fn=_unlimited_stack, check=True)
    fluids = sorted(f for f in os.listdir(wd) if re.fullmatch(r"fluid\d\d", f))
    out = {"input": f"tests/golden/main_inputs/{name}.input",
           "program": "oracle/_ref/g3m3_main (oracle/Makefile target main)",
           "grid": [NX, NY, NZ, NGRD], "fluid": {}}
    for f in fluids:
        print(f"  digesting {f}", flush=True)
        out["fluid"][f] = digest_fluid(os.path.join(wd, f))
        os.remove(os.path.join(wd, f))
    gdir = os.path.join(wd, "figures", "data")
    graphs = sorted(f for f in os.listdir(gdir) if f.endswith(".dat"))
    if graphs:
        out["graphing"] = {}
        for f in graphs:
            out["graphing"][f] = digest_graph(os.path.join(gdir, f))
            os.remove(os.path.join(gdir, f))
    out["steps"] = parse_speeds(os.path.join(wd, "speeds.dat"))
    out["fluxes"] = parse_fluxes(os.path.join(wd, "fluxes.dat"))
    return out


def devmath(name: str) -> None:
    sys.path[:0] = [os.path.dirname(HERE), ROOT]
    from mainpin import digest_records, oracle_main_loop
    out = {"input": f"tests/golden/main_inputs/{name}.input",
           "program": "the oracle (tests/mainpin.py:oracle_main_loop), powf/expf in double",
           "grid": [NX, NY, NZ, NGRD], "fluid": {}}
    log = None
    for fname, t, recs, log in oracle_main_loop(name, transcendentals="double"):
        print(f"  {fname}: t = {t!r}", flush=True)
        out["fluid"][fname] = digest_records(recs, (NX, NY, NZ, NGRD))
    out["steps"] = [list(r) for r in log["steps"]]
    fl = []
    for ut, box, qi, hi, oi, qo, ho, oo in log["fluxes"]:
        fl.append([ut, box, qi, hi, oi, "grid_in"])
        fl.append([ut, box, qo, ho, oo, "grid_out"])
    out["fluxes"] = fl
    dst = os.path.join(HERE, f"main_{name}_devmath.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {dst}")


def main(names=INPUTS):
    if not os.access(BIN, os.X_OK):
        sys.exit(f"{BIN} missing: run `make -C oracle main` first")
    for name in names:
        res = run_one(name)
        dst = os.path.join(HERE, f"main_{name}.json")
        with open(dst, "w") as f:
            json.dump(res, f, indent=1)
        print(f"wrote {dst}: {len(res['fluid'])} fluid files, {len(res['steps'])} box-steps")


if __name__ == "__main__":
    if sys.argv[1:2] == ["--devmath"]:
        for n in sys.argv[2:]:
            devmath(n)
    else:
        main(sys.argv[1:] or INPUTS)
