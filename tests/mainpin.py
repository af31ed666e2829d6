"""Comparison of a run against what the reference's whole program wrote
(tests/golden/main_*.json, made by tests/golden/make_main_golden.py): fluid
records, box-step times (speeds.dat) and wall fluxes (fluxes.dat)."""
from __future__ import annotations

import base64
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
FIELD_RECORDS = ["qrho", "qpx", "qpy", "qpz", "qpresx", "qpresy", "qpresz", "qpresxy",
                 "qpresxz", "qpresyz", "hrho", "hpx", "hpy", "hpz", "hpresx", "hpresy",
                 "hpresz", "hpresxy", "hpresxz", "hpresyz", "orho", "opx", "opy", "opz",
                 "opresx", "opresy", "opresz", "opresxy", "opresxz", "opresyz", "bx", "by",
                 "bz", "epres", "bx0", "by0", "bz0"]
LIST_PARTS = ("parm_srf", "parm_mid", "parm_zero", "ijzero", "numzero", "ijmid", "nummid",
              "ijsrf", "numsrf")


def golden(name: str) -> dict:
    with open(os.path.join(GOLDEN, f"main_{name}.json")) as f:
        return json.load(f)


def input_path(name: str) -> str:
    return os.path.join(GOLDEN, "main_inputs", name + ".input")


def sample_index(nx, ny, nz, ngrd) -> np.ndarray:
    """The sample cells of tests/golden/make_main_golden.py."""
    V = nx * ny * nz
    idx = []
    for b in range(2):
        base = b * V + (nz // 2) * nx * ny
        idx += [base + (ny // 2) * nx + i for i in range(nx)]
        idx += [base + j * nx + nx // 2 for j in range(ny)]
    rng = np.random.default_rng(20261207)
    idx += list(rng.integers(0, ngrd * V, 256))
    return np.array(idx, np.int64)


def records_of(t: float, arrays: dict, lists: dict) -> list:
    """Payload bytes of the 39 fluid-file records (code/3d_main.f90:3312-3349):
    t, the 37 arrays in record order, then the lists record."""
    out = [np.float64(t).tobytes()]
    out += [np.ascontiguousarray(arrays[k], np.float32).tobytes() for k in FIELD_RECORDS]
    out.append(b"".join(np.ascontiguousarray(lists[k]).tobytes() for k in LIST_PARTS))
    return out


def read_fluid_records(path: str) -> list:
    out = []
    with open(path, "rb") as f:
        while True:
            head = f.read(4)
            if not head:
                break
            n = int(np.frombuffer(head, "<i4")[0])
            data = f.read(n)
            assert int(np.frombuffer(f.read(4), "<i4")[0]) == n
            out.append(data)
    return out


def compare_records(recs: list, gold: dict, dims, exact: bool, tol: float = 0.0,
                    tol_sumsq: float | None = None) -> list:
    """Differences between record payloads and a golden fluid file: exact
    compares SHA-256 digests (bit for bit); otherwise each field's samples
    within ``tol`` of the field's largest golden sample magnitude and its
    sum of squares within ``tol_sumsq`` (default ``tol``) relative, t bit
    for bit and the lists record not compared.  Returns a list of
    problems."""
    tol_ss = tol if tol_sumsq is None else tol_sumsq
    bad = []
    rows = gold["records"]
    if len(recs) != len(rows):
        return [f"{len(recs)} records, golden {len(rows)}"]
    idx = sample_index(*dims)
    for data, row in zip(recs, rows):
        name = row["name"]
        if len(data) != row["bytes"]:
            bad.append(f"{name}: {len(data)} bytes, golden {row['bytes']}")
            continue
        if not exact and name == "lists":
            # the saved boundary values are initial densities and pressures,
            # so they carry whatever the fields carry: checked bit for bit
            # only (exact=True)
            continue
        if exact or name == "t":
            if hashlib.sha256(data).hexdigest() != row["sha256"]:
                if name in ("t", "lists") or "samples" not in row:
                    bad.append(f"{name}: digest differs")
                else:
                    a = np.frombuffer(data, "<f4")
                    want = np.frombuffer(base64.b64decode(row["samples"]), "<f4")
                    got = a[idx]
                    nd = int((got.view(np.uint32) != want.view(np.uint32)).sum())
                    bad.append(f"{name}: digest differs ({nd} of {len(idx)} samples differ, "
                               f"max |d| {float(np.abs(got.astype(float) - want).max()):.3e})")
            continue
        a = np.frombuffer(data, "<f4").astype(np.float64)
        want = np.frombuffer(base64.b64decode(row["samples"]), "<f4").astype(np.float64)
        scale = max(float(np.abs(want).max()), 1e-30)
        d = float(np.abs(a[idx] - want).max()) / scale
        ss = float((a * a).sum())
        dss = abs(ss - row["sumsq"]) / max(abs(row["sumsq"]), 1e-30)
        if not (d <= tol and dss <= tol_ss):
            bad.append(f"{name}: samples {d:.3e}, sum of squares {dss:.3e} "
                       f"(bars {tol}, {tol_ss})")
    return bad


def compare_steps(log: list, gold: dict, rel: float = 0.0) -> list:
    """(box, t, t_old, t_new) records against speeds.dat: the same boxes in
    the same order, times exactly (rel = 0) or within rel."""
    want = gold["steps"]
    if len(log) != len(want):
        return [f"{len(log)} box-steps, the reference took {len(want)}"]
    bad = []
    for q, (g, w) in enumerate(zip(log, want)):
        off = [abs(float(x) - float(y)) > rel * abs(float(y)) for x, y in zip(g[1:], w[1:])]
        if int(g[0]) != w[0] or any(off):
            bad.append(f"box-step {q}: {tuple(g)} vs {tuple(w)}")
    return bad[:10]


def compare_fluxes(log: list, gold: dict, rel: float = 0.0) -> list:
    """(ut, box, q/h/o in, q/h/o out) per box and step against fluxes.dat:
    ut exactly, the single-precision fluxes exactly (rel = 0) or within rel."""
    want = gold["fluxes"]
    if 2 * len(log) != len(want):
        return [f"{len(log)} flux rows, the reference wrote {len(want) // 2} pairs"]
    bad = []
    for q, row in enumerate(log):
        ut, box, qi, hi, oi, qo, ho, oo = row
        w_in, w_out = want[2 * q], want[2 * q + 1]
        if w_in[5] != "grid_in" or w_out[5] != "grid_out" or w_in[1] != box:
            bad.append(f"row {q}: layout")
            continue
        if float(ut) != w_in[0]:
            bad.append(f"row {q}: ut {ut!r} vs {w_in[0]!r}")
        for got, ref in zip((qi, hi, oi, qo, ho, oo), (*w_in[2:5], *w_out[2:5])):
            g32, r32 = np.float32(got), np.float32(ref)
            if rel == 0.0:
                if g32 != r32:
                    bad.append(f"row {q} box {box}: flux {g32!r} vs {r32!r}")
            elif abs(float(g32) - float(r32)) > rel * max(abs(float(r32)), 1e-30):
                bad.append(f"row {q} box {box}: flux {g32!r} vs {r32!r}")
    return bad[:10]


def oracle_main_loop(name: str, nsteps: int | None = None, transcendentals: str = "libm"):
    """The main program on the CPU: g3m3_amd.mainprog's host set-up, the
    oracle's initialisation block, OracleRun driving the reference's
    compiled subroutines, the oracle's torus source, and the loop's order
    (code/3d_main.f90:1925-3361: wall fluxes, full step, torus, fluid file).
    Yields, per fluid file, (name, t, records); ``log`` collects the
    box-step and flux records."""
    from g3m3_amd import mainprog as M
    from oracle import ref
    from oracle import restate as O
    from oracle.boxstep import OracleRun
    O.set_transcendentals(transcendentals)
    try:
        S = M.setup(M.read_input(input_path(name)))
        wl = S.wl
        L, g = wl.lists, wl.grid
        gmin = np.ascontiguousarray(np.stack([g["grd_xmin"], g["grd_ymin"], g["grd_zmin"]], 1))
        gmax = np.ascontiguousarray(np.stack([g["grd_xmax"], g["grd_ymax"], g["grd_zmax"]], 1))
        O.init_plasma(wl.fields, wl.fields["bx0"], wl.fields["by0"], wl.fields["bz0"],
                      wl.fields["rst"], wl.nx, wl.ny, wl.nz, wl.ngrd, int(wl.scal["mbndry"]),
                      *S.caps, gmin, gmax, L["parm_srf"], L["parm_mid"], L["parm_zero"],
                      L["ijsrf"], L["ijmid"], L["ijzero"], L["numsrf"], L["nummid"], L["numzero"],
                      S.init)
        o = OracleRun(wl, "ref")
        o.t = S.t
        ref.with_big_stack(o.init_dt)
        log = {"steps": [], "fluxes": [], "torus": []}
        tinj, ts1, tmax = S.tinj, S.ts1, S.tmax
        fluid_no, steps = 1, 0
        while o.t < float(tmax) and (nsteps is None or steps < nsteps):
            rows = []
            for box in range(o.ngrd, 0, -1):
                out = [np.zeros(1, np.float32) for _ in range(6)]
                O.flux_counter(*[o.a[k] for k in ("qpx", "qpy", "qpz", "hpx", "hpy", "hpz",
                                                  "opx", "opy", "opz")], o.nx, o.ny, o.nz,
                               o.ngrd, box, o.s["rmassq"], o.s["rmassh"], o.s["rmasso"],
                               S.flux_scale[box - 1], *out)
                qi, qo, hi, ho, oi, oo = (float(x[0]) for x in out)
                rows.append((box, qi, hi, oi, qo, ho, oo))
            n0 = len(o.step_log)
            ref.with_big_stack(o.full_step)
            steps += 1
            log["fluxes"] += [(o.ut, *r) for r in rows]
            log["steps"] += [(b, o.t, t0, t1) for b, t0, t1 in o.step_log[n0:]]
            if o.t >= float(tinj) and S.inp.get("ringo", False):
                tot = O.torus_inject({k: o.a[k] for k in o.a}, o.a["bx0"], o.a["by0"],
                                     o.a["bz0"], o.nx, o.ny, o.nz, gmin, gmax, S.torus)
                log["torus"].append((o.t, *[float(x) for x in tot]))
                tinj = np.float32(tinj + S.deltinj)
            if o.t >= float(ts1):
                yield f"fluid{fluid_no:02d}", o.t, records_of(o.t, o.a, wl.lists), log
                fluid_no += 1
                ts1 = np.float32(ts1 + np.float32(S.inp.get("tsave", 0.0)))
    finally:
        O.set_transcendentals("libm")


def digest_records(recs, dims) -> dict:
    """The golden-file digest of record payloads (make_main_golden.py's)."""
    idx = sample_index(*dims)
    rows = []
    names = ["t"] + FIELD_RECORDS + ["lists"]
    for name, data in zip(names, recs):
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


def compare_graphs(data_dir: str, gold: dict, rel: float = 2e-6) -> dict:
    """Plotting tables under data_dir against the reference's
    (tests/golden/main_pin_graph.json): the same files, line counts and
    header lines, and at the sampled rows every printed value within rel
    of the reference's (the device's log10 and a few others may move the
    7th printed digit).  Returns counts and the problems found."""
    import re
    import zlib
    want = gold["graphing"]
    have = sorted(f for f in os.listdir(data_dir) if f.endswith(".dat"))
    out = {"files": len(have), "values": 0, "text_equal": 0, "problems": []}
    if have != sorted(want):
        out["problems"].append(f"files differ: {sorted(set(have) ^ set(want))[:5]}")
        return out
    for name in have:
        w = want[name]
        lines = open(os.path.join(data_dir, name)).read().split("\n")
        if lines and lines[-1] == "":
            lines.pop()
        if len(lines) != w["lines"]:
            out["problems"].append(f"{name}: {len(lines)} lines, reference {w['lines']}")
            continue
        if lines[:len(w["head"])] != w["head"]:
            out["problems"].append(f"{name}: header differs")
        rows = json.loads(zlib.decompress(base64.b64decode(w["rows"])))
        head = len(w["head"])
        for r, text in rows.items():
            got = lines[head + int(r)]
            wt, gt = text.split(), got.split()
            if len(wt) != len(gt):
                out["problems"].append(f"{name} row {r}: column count")
                continue
            for a, b in zip(gt, wt):
                out["values"] += 1
                if a == b:
                    out["text_equal"] += 1
                    continue
                x, y = float(a), float(b)
                if not (abs(x - y) <= rel * abs(y) or (np.isnan(x) and np.isnan(y))):
                    out["problems"].append(f"{name} row {r}: {a} vs {b}")
    return out
