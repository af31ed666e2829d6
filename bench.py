#!/usr/bin/env python3
"""Benchmark of the MI355X G3M^3 hot path: full steps of the nested
121x121x61 workload, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

A "step" is one full step of the reference's outer loop
(code/3d_main.f90:1925-3095): dt synchronisation, every box-step the
subcycling schedule asks for (31 for five levels at a CFL ratio of 2),
nested-grid sync.  ``value`` = cell-updates/s = sum over executed
box-steps of nx*ny*nz, divided by the wall time of the K timed steps.
The divB clean-up (off-step in the reference, at checkpoints) is timed
separately and reported in ``divb``.

One JSON line is printed.  ``roofline`` is for the kernel named by
--prof-kernel (timed with HIP events inside the library over the timed
region); ``cpu_baseline`` is the reference's own OpenMP build timed on this
host over one full step of the same workload (or, where that build is
absent, the NumPy restatement on one box-step: kind "port").
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# compulsory FP32 traffic of one cell-update: the four stages separated by
# grid-wide dependencies (predictor 71+33, corrector 71+34, Lapidus 34+34,
# FCT 68+34 accesses of 4 bytes)
WHOLE_STEP_BYTES = 1384
METRIC = "cell-updates/s, full step, 5 nested 121x121x61 grids; adjoint/forward cost"

# Algorithmic HBM bytes per unit of each instrumented kernel (unit: one
# interior cell, one field for the FCT kernels), counting each input and
# output array once -- what the kernel must move if nothing is re-read.
KERNEL_BYTES = {
    # algorithmic bytes per unit (each input and output array counted once)
    # native runtime:
    # the push stage: per cell the compulsory FP32 traffic of the two
    # Lax-Wendroff stages it completes, averaged over its two calls per
    # box-step -- predictor: state 34 + wrkepres + b0 3 read, 33 written
    # (284 B); corrector: wrk 34 + old 34 + b0 3 read, 34 written (420 B).
    # Scratch (E, ev, tv, bs) is overhead and not counted.
    "k_push": ((284 + 420) // 2, "all"),
    # per cell: 30 species fields + epres + b 3 + b0 3 read (maxima pass)
    "k_validate": (37 * 4, "all"),
    # per interior cell: 34 fields read, 34 written
    "k_lap": (68 * 4, "interior"),
    # per cell and field: wrk, old read, out written (P stays in LDS)
    "k_fct_fused": (3 * 4, "all_x_fields"),
    # per cell: b 3, b0 3, 9 momenta, 3 densities, epres read; E, ev, tv, bs written
    "k_efield": (31 * 4, "all"),
    # per shell point and field: target read and written, and the coarse
    # cells under the shell once (about 0.8 per point: two time levels)
    "k_flanks_table": (3 * 4, "shell_x_fields"),
    # per-subroutine path:
    # wrk 10 + old 10 + qv 3 + bs 3 + btot + E 3 + tv 3 reads, 10 writes
    "k_push_ion": (43 * 4, "interior"),
    # v 3 + 10 fields read, 10 written
    "k_lapidus": (23 * 4, "interior"),
    # 30 species fields read
    "k_speed_flag": (30 * 4, "interior"),
    "k_copy_fields": (2 * 4, "all_x_fields"),
}


def units_per_launch(kind, wl, nfields=34):
    nx, ny, nz = wl.nx, wl.ny, wl.nz
    if kind == "interior":
        return (nx - 2) * (ny - 2) * (nz - 2)
    if kind == "all":
        return nx * ny * nz
    if kind == "shell_x_fields":   # the 2-layer shell of a box
        return (4 * nx * ny + 4 * nx * (nz - 4) + 4 * (ny - 4) * (nz - 4)) * nfields
    return nx * ny * nz * nfields


# The kernels one call of an instrumented stage launches (the library times
# the stage as one unit): the k-marching species kernels plus the per-cell
# kernel for B and the electrons.
STAGE_KERNELS = {"k_push": ("k_push_march", "k_push"), "k_lap": ("k_lap_march", "k_lap")}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per call of stage `kernel` (summed over the kernels it
    launches) measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counters,
    calibrated on a plain copy (the commands and the calibration are in
    profiles/traffic_native_<workload>.md)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"traffic_native_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        data = json.load(fh)
    total, found = 0, False
    for base in STAGE_KERNELS.get(kernel, (kernel,)):
        for name, rec in data.get("kernels", {}).items():
            if name == base or name.startswith(base + "<"):
                total += rec["traffic_bytes"]
                found = True
    return (total, f"profiles/traffic_native_{workload}.md") if found else (None, None)


def batch_case(n: int = 32, steps: int = 20, streams: int = 8):
    """Throughput of n independent runs stepped one after another (each run
    alone on the GPU) and stepped together (g3m3_amd.driver.Batch)."""
    import torch
    from g3m3_amd.driver import Batch, FastRun
    from g3m3_amd.workload import make_batch_workloads
    wls = make_batch_workloads(n)
    runs = [FastRun(wl) for wl in wls]
    for r in runs:
        r.init_dt()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in runs:
        for _ in range(steps):
            r.full_step()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    cu_plain = sum(r.cell_updates for r in runs)
    del runs
    b = Batch(wls, streams=streams)
    b.init_dt()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.full_step()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    cu_batch = b.cell_updates
    del b
    # the same runs replaying captured step graphs, time step on the device
    g = Batch(wls, streams=streams)
    g.init_dt()
    g.graph_steps(0)            # capture and instantiate, outside the timing
    torch.cuda.synchronize()
    cu0 = g.cell_updates
    t0 = time.perf_counter()
    g.graph_steps(steps)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    cu_graph = g.cell_updates - cu0
    del g
    # the same runs in lockstep groups: one multi-run launch per stage and
    # group, one captured graph per step and group, the groups on their own
    # streams (set-up and recording outside the timing); every group count
    # tried, the fastest reported
    fused = {}
    for ng in (1, 2, 4, 8):
        fb = Batch(wls, streams=1)
        fb.init_dt()
        fb.fused_steps(0, groups=ng)
        torch.cuda.synchronize()
        cu0 = fb.cell_updates
        t0 = time.perf_counter()
        fb.fused_steps(steps, groups=ng)
        torch.cuda.synchronize()
        fused[ng] = ((fb.cell_updates - cu0) / (time.perf_counter() - t0), fb.fused_launches)
        del fb
    best = max(fused, key=lambda g: fused[g][0])
    plain = cu_plain / t_plain
    return {"runs": n, "steps": steps, "streams": streams,
            "cells_min_max": [min(w.cells_per_box for w in wls), max(w.cells_per_box for w in wls)],
            "plain_cell_updates_per_s": plain,
            "batched_cell_updates_per_s": cu_batch / t_batch,
            "speedup": (cu_batch / t_batch) / plain,
            "graph_cell_updates_per_s": cu_graph / t_graph,
            "graph_speedup": (cu_graph / t_graph) / plain,
            "fused_cell_updates_per_s": fused[best][0],
            "fused_speedup": fused[best][0] / plain,
            "fused_groups": best,
            "fused_launches_per_step": fused[best][1],
            "fused_candidates": {str(g): {"cell_updates_per_s": v[0], "launches_per_step": v[1]}
                                 for g, v in fused.items()}}


def gradient_case():
    """The differentiation configuration (three nested 81x81x41 boxes, 50 full
    steps, 8 craft sampled every 10 steps) against a target run with
    perturbed inputs: the loss, its gradient by central differences, and
    the device time of one forward run and of the gradient's 12."""
    from g3m3_amd import misfit as M
    cfg = M.MisfitConfig()
    theta = M.THETA0.copy()
    target_theta = np.array([1.0, -1.0, -6.0, 5.5, 420.0, 1e-3])
    target, _ = M.forward(target_theta, cfg)
    s, fwd = M.forward(theta, cfg)
    runs = []
    g, secs = M.gradient_fd(theta, target, cfg, samples_out=runs)
    g1, secs1 = M.gradient_fd(theta, target, cfg, together=False)
    assert np.array_equal(g, g1)
    # parameters whose two runs sample bit-identical values at every craft:
    # their change has not reached any craft within the run
    unreached = [M.THETA_NAMES[i] for i in range(len(theta))
                 if np.array_equal(np.asarray(runs[2 * i]).view(np.uint32),
                                   np.asarray(runs[2 * i + 1]).view(np.uint32))]
    return {"config": "3 nested 81x81x41 boxes, 50 full steps, 8 craft sampled every 10 steps",
            "theta": dict(zip(M.THETA_NAMES, theta.tolist())),
            "target_theta": dict(zip(M.THETA_NAMES, target_theta.tolist())),
            "loss": M.loss(s, target), "gradient": dict(zip(M.THETA_NAMES, g.tolist())),
            "method": "central differences, h = 1e-3 |theta_i|, two forward runs per "
                      "parameter stepped together on 8 streams (no adjoint)",
            "forward_s": round(fwd, 4), "gradient_s": round(secs, 4),
            "gradient_over_forward": round(secs / fwd, 2),
            "gradient_runs_one_at_a_time_s": round(secs1, 4),
            "unreached": unreached,
            "unreached_note": "a listed parameter's +h and -h runs give bit-identical samples "
                              "at every craft, so its difference quotient is exactly 0: the "
                              "upstream IMF is imposed on the inflow face of the outer box, "
                              "and in 50 steps the change has not been carried to any craft"}


def adjoint_case(steps: int = 50, every: int = 10, ckpt: int = 10, ncraft: int = 8):
    """The checkpointed adjoint of the field update (g3m3_amd.adjoint) on one
    121x121x61 box upstream of the planet: a craft misfit over ``steps``
    steps against a target run with other controls plus 5% noise, its gradient in
    (eta, sbx, sby, sbz) and in the initial field, and the wall time of the
    forward rollout and of the backward pass (recompute from checkpoints
    every ``ckpt`` steps + adjoint kernels)."""
    import torch
    from g3m3_amd import adjoint as A
    s = A.field_setup(nx=121, ny=121, nz=61, spacing=0.5)
    m = A.FieldModel(s)
    pos = torch.from_numpy(A.craft_positions(s, ncraft, seed=12)).cuda()
    ctl_true = torch.tensor([1e-3, 0.04, 0.02, -0.20], dtype=torch.float64)
    ctl0 = torch.tensor([2e-3, 0.0, 0.0, -0.24], dtype=torch.float64)
    with torch.no_grad():
        target = A.rollout(ctl_true, m.new_state(), m, steps, every, pos).clone()
    # observation noise: 5% of the samples' rms (seeded)
    gen = torch.Generator().manual_seed(13)
    rms = float(target.pow(2).mean().sqrt())
    target += (0.05 * rms) * torch.randn(target.shape, generator=gen).cuda()

    def grad():
        c = ctl0.clone().requires_grad_(True)
        b0 = m.new_state().requires_grad_(True)
        out = A.rollout(c, b0, m, steps, every, pos, ckpt)
        L = ((out - target) ** 2).sum()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        return float(L.detach()), c.grad.numpy(), time.perf_counter() - t0

    grad()                                   # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        A.rollout(ctl0, m.new_state(), m, steps, every, pos)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    L, g, bwd = grad()
    m.close()
    return {"config": f"one 121x121x61 box, frozen plasma, {steps} steps, {ncraft} craft "
                      f"sampling b every {every} steps, checkpoints every {ckpt} steps",
            "controls": dict(zip(A.CTL_NAMES, ctl0.tolist())),
            "target_controls": dict(zip(A.CTL_NAMES, ctl_true.tolist())),
            "loss": L, "gradient": dict(zip(A.CTL_NAMES, g.tolist())),
            "forward_ms": round(fwd * 1e3, 2), "backward_ms": round(bwd * 1e3, 2),
            "backward_over_forward": round(bwd / fwd, 2),
            "method": "hand-written adjoint kernels, recompute from checkpoints"}


def full_adjoint_case(steps: int = 50, every: int = 10, ckpt: int = 10, ncraft: int = 8):
    """The checkpointed adjoint of the whole box-step (g3m3_amd.adjoint
    full model) on one 121x121x61 box upstream of the planet: a misfit of
    density, momentum, electron pressure and field at ``ncraft`` craft near
    the inflow face over ``steps`` steps, against a run with other upstream
    values and resistivity plus 5% noise; its gradient in the nine controls
    and the initial state, and the wall time of the forward rollout and of
    the backward pass (recompute from checkpoints + adjoint kernels)."""
    import torch
    from g3m3_amd import adjoint as A
    s = A.full_setup(nx=121, ny=121, nz=61, spacing=0.5)
    m = A.FullModel(s)
    pos_h = A.craft_positions(s, ncraft, seed=14)
    h = float(s.grid["xspac"][0])
    x0 = float(s.grid["grid_minvals"][0][0])
    pos_h[:, 0] = (x0 + h * np.random.default_rng(5).uniform(2.0, 12.0, ncraft)).astype(np.float32)
    pos = torch.from_numpy(pos_h).cuda()
    ctl0 = torch.tensor(s.ctl.astype(np.float64))
    ctl_true = ctl0.clone()
    ctl_true[0], ctl_true[3] = 5e-4, -0.20
    ctl_true[4] *= 1.05
    ctl_true[5] *= 0.97
    with torch.no_grad():
        target = A.full_rollout(ctl_true, m.new_state(), m, steps, every, pos).clone()
    gen = torch.Generator().manual_seed(15)
    rms = target.pow(2).mean(dim=(0, 1)).sqrt()
    target += 0.05 * rms * torch.randn(target.shape, generator=gen).cuda()
    wt = 1.0 / rms

    def grad():
        c = ctl0.clone().requires_grad_(True)
        s0 = m.new_state().requires_grad_(True)
        out = A.full_rollout(c, s0, m, steps, every, pos, ckpt)
        L = (((out - target) * wt) ** 2).sum()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        return float(L.detach()), c.grad.numpy(), time.perf_counter() - t0

    grad()                                   # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        A.full_rollout(ctl0, m.new_state(), m, steps, every, pos)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    L, g, bwd = grad()
    m.close()
    return {"config": f"one 121x121x61 box, every field evolving, {steps} steps, {ncraft} craft "
                      f"sampling 8 quantities every {every} steps, checkpoints every {ckpt} steps",
            "controls": dict(zip(A.FULL_CTL_NAMES, ctl0.tolist())),
            "target_controls": dict(zip(A.FULL_CTL_NAMES, ctl_true.tolist())),
            "loss": L, "gradient": dict(zip(A.FULL_CTL_NAMES, g.tolist())),
            "forward_ms": round(fwd * 1e3, 2), "backward_ms": round(bwd * 1e3, 2),
            "backward_over_forward": round(bwd / fwd, 2),
            "method": "hand-written adjoint kernels, recompute from checkpoints"}


def nested_adjoint_case(nbox: int = 5, steps: int = 4, every: int = 2, ckpt: int = 2,
                        ncraft: int = 8):
    """The checkpointed adjoint of the whole box-step through ``nbox`` nested
    121x121x61 boxes (g3m3_amd.adjoint_nested; inner boxes bounded by their
    parent's flanks and restricted into its core, on the subcycling
    schedule): a misfit of density, momentum, electron pressure and field at
    ``ncraft`` craft, half inside the innermost box and half near the outer
    box's inflow face, over ``steps`` full steps; its gradient in the nine
    controls and every box's initial state, and the wall time of the forward
    rollout and of the backward pass."""
    import torch
    from g3m3_amd import adjoint as A
    from g3m3_amd import adjoint_nested as N
    ss = N.nested_setup(nbox, nx=121, ny=121, nz=61, spacing=1.0)
    m = N.NestedFullModel(ss)
    outer = ss[-1]
    half = ncraft // 2
    p_in = A.craft_positions(ss[0], half, seed=31, margin=3.0)
    p_out = A.craft_positions(outer, ncraft - half, seed=32, margin=3.0)
    h = float(outer.grid["xspac"][0])
    x0 = float(outer.grid["grid_minvals"][0][0])
    p_out[:, 0] = (x0 + h * np.random.default_rng(33).uniform(2.0, 12.0, ncraft - half)
                   ).astype(np.float32)
    pos = torch.from_numpy(np.ascontiguousarray(np.concatenate([p_in, p_out]))).cuda()
    ctl0 = torch.tensor(outer.ctl.astype(np.float64))
    ctl_true = ctl0.clone()
    ctl_true[0], ctl_true[3] = 5e-4, -0.20
    ctl_true[4] *= 1.05
    ctl_true[5] *= 0.97
    with torch.no_grad():
        target = N.nested_rollout(ctl_true, m.new_states(), m, steps, every, pos).clone()
    gen = torch.Generator().manual_seed(34)
    rms = target.pow(2).mean(dim=(0, 1)).sqrt()
    target += 0.05 * rms * torch.randn(target.shape, generator=gen).cuda()
    wt = 1.0 / rms

    def grad():
        c = ctl0.clone().requires_grad_(True)
        s0 = [s.requires_grad_(True) for s in m.new_states()]
        out = N.nested_rollout(c, s0, m, steps, every, pos, ckpt)
        L = (((out - target) * wt) ** 2).sum()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        return float(L.detach()), c.grad.numpy(), time.perf_counter() - t0

    grad()                                   # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        N.nested_rollout(ctl0, m.new_states(), m, steps, every, pos)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    L, g, bwd = grad()
    nbs = m.box_steps
    m.close()
    return {"config": f"{nbox} nested 121x121x61 boxes, spacing "
                      f"{[float(s.scal['rx']) for s in ss]}, every field evolving, {steps} full "
                      f"steps ({nbs} box-steps each), {ncraft} craft sampling 8 quantities "
                      f"every {every} full steps, checkpoints every {ckpt} full steps",
            "controls": dict(zip(A.FULL_CTL_NAMES, ctl0.tolist())),
            "target_controls": dict(zip(A.FULL_CTL_NAMES, ctl_true.tolist())),
            "loss": L, "gradient": dict(zip(A.FULL_CTL_NAMES, g.tolist())),
            "forward_ms": round(fwd * 1e3, 2), "backward_ms": round(bwd * 1e3, 2),
            "backward_over_forward": round(bwd / fwd, 2),
            "method": "hand-written adjoint kernels (box-steps, flanks, corer, sampling), "
                      "recompute from checkpoints"}


def precision_case(fast_mode: int):
    """divB on a 241x241x121 dipole + 1% noise field to rel-residual 1e-8 with
    FP32 vectors, FP64 dots and FP64 residual refinement (mode 3); the
    FP32-vector CG alone (mode 2) at 1e-6 beside it."""
    import ctypes
    import torch
    from g3m3_amd import _lib
    from g3m3_amd.workload import precision_field
    host = precision_field()
    nz, ny, nx = host[0].shape[1:]
    npts = nx * ny * nz
    xs = np.array([1.0], np.float32)
    # |b|^2 of the system's right side (central-difference divergence / 2)
    X, Y, Z = (h[0].astype(np.float64) for h in host)
    rhs = ((X[1:-1, 1:-1, 2:] - X[1:-1, 1:-1, :-2]) + (Y[1:-1, 2:, 1:-1] - Y[1:-1, :-2, 1:-1])
           + (Z[2:, 1:-1, 1:-1] - Z[:-2, 1:-1, 1:-1])) / 2.0
    b2 = float((rhs * rhs).sum())
    del X, Y, Z, rhs
    res = {"grid": f"{nx}x{ny}x{nz}"}
    for mode, rel in ((2, 1e-6), (3, 1e-8)):
        dev = [torch.from_numpy(h).cuda() for h in host]
        torch.cuda.synchronize()
        it = ctypes.c_longlong(0)
        r2 = ctypes.c_double(0)
        t0 = time.perf_counter()
        st = _lib.call("divb_correct_ex", nx, ny, nz, 1, 1, xs, np.float32(rel / np.sqrt(npts)),
                       dev[0], dev[1], dev[2], mode, ctypes.c_longlong(0), it, r2)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _lib.check("divb precision")
        res[f"mode{mode}"] = {"rel_residual_target": rel, "status": st, "iterations": it.value,
                              "seconds": round(dt, 3),
                              "achieved_rel_residual": float(np.sqrt(max(r2.value, 0.0) / b2))}
    return res


def _cpu_child(workload: str, threads: int, what: str) -> dict:
    """Runs in a child process (no GPU): the reference's OpenMP build over
    one full step (what = "step") or one box-step of the finest box (what =
    "box") of the workload, timed."""
    from g3m3_amd.workload import make_workload
    from oracle import ref
    from oracle.boxstep import OracleRun
    wl = make_workload(workload)
    out = {}

    def go():
        r = OracleRun(wl, "ref_omp", threads=threads)
        r.init_dt()
        if what == "step":
            t0 = time.perf_counter()
            r.full_step()
        else:
            r.t_step[:] = r.t_stepnew
            r.t_step[:] = np.minimum.accumulate(r.t_step[::-1])[::-1]
            t0 = time.perf_counter()
            r.box_step(1)
        out.update(cu=r.cell_updates, dt=time.perf_counter() - t0, bs=r.box_steps)
    ref.with_big_stack(go)
    return out


def cpu_baseline(workload: str, threads: int):
    """The reference's OpenMP build (flang -O3 -march=x86-64-v3 -fopenmp) on
    this host: one full step of the same workload at `threads` threads, and
    one box-step at 1 thread beside it.  Each runs in a child process, so a
    failure there cannot end the benchmark; where that build is absent, the
    NumPy restatement on one box-step (kind "port")."""
    import subprocess
    from oracle import ref
    if not ref.available(omp=True):
        from g3m3_amd.workload import make_workload
        from oracle.boxstep import OracleRun
        wl = make_workload(workload)
        r = OracleRun(wl, "restate")
        r.init_dt()
        r.t_step[:] = r.t_stepnew
        r.t_step[:] = np.minimum.accumulate(r.t_step[::-1])[::-1]
        t0 = time.perf_counter()
        r.box_step(wl.ngrd)
        dt = time.perf_counter() - t0
        return {"value": r.cell_updates / dt, "unit": "cell-updates/s", "cores": 1,
                "kind": "port", "sample": f"1 box-step of the outermost box of {workload} by "
                                          f"the NumPy restatement, {dt:.1f} s"}
    res = {}
    for label, th, what in (("multi", threads, "step"), ("single", 1, "box")):
        env = dict(os.environ, OMP_NUM_THREADS=str(th), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", workload, str(th), what]
        try:
            p = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=240)
            res[label] = json.loads(p.stdout.strip().splitlines()[-1]) if p.returncode == 0 \
                else {"error": f"exit {p.returncode}: {p.stderr[-300:]}"}
        except Exception as e:   # a failed baseline is reported, never fatal
            res[label] = {"error": str(e)}
    m, s1 = res["multi"], res["single"]
    if "error" in m:
        return {"value": None, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
                "error": m["error"]}
    out = {"value": m["cu"] / m["dt"], "unit": "cell-updates/s", "cores": threads,
           "kind": "reference",
           "sample": f"1 full step ({m['bs']} box-steps, {m['cu']} cell-updates) of {workload}, "
                     f"the reference's sub_*.f90 built with flang -O3 -march=x86-64-v3 "
                     f"-fopenmp, {threads} threads, {m['dt']:.1f} s"}
    if "error" not in s1:
        out["single_thread"] = {
            "value": s1["cu"] / s1["dt"], "cores": 1,
            "sample": f"1 box-step of box 1 ({s1['cu']} cell-updates), 1 thread, "
                      f"{s1['dt']:.1f} s"}
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv[:1] == ["--cpu-child"]:
        print(json.dumps(_cpu_child(argv[1], int(argv[2]), argv[3])))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="nested121")
    ap.add_argument("--prof-kernel", default="k_push")
    ap.add_argument("--driver", choices=("native", "persub"), default="native",
                    help="native: fused kernels issued by the C++ runtime; persub: one C-ABI "
                         "call per reference subroutine")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-divb", action="store_true")
    ap.add_argument("--divb-mode", type=int, default=2,
                    help="0/1: the reference's CGNR (single / double scalars); 2: conjugate "
                         "gradients on the same system (FP64 dots)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the divB precision case (241x241x121 dipole + 1%% noise, "
                         "mode 3, FP64 residual refinement, to rel-residual 1e-8)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batching case (32 independent single-grid runs, "
                         "41x41x21 .. 121x121x61, 20 full steps each, one at a time, together "
                         "on 8 streams, and replayed from captured graphs)")
    ap.add_argument("--species-kernels", choices=("auto", "march", "cell"), default="auto",
                    help="native driver: the species push / Lapidus kernels (auto: the run's "
                         "autotuner picks; counter runs fix the family the timed runs pick, "
                         "since counter collection changes the timings the tuner sees)")
    ap.add_argument("--no-gradient", action="store_true",
                    help="skip the differentiation case (3 nested 81x81x41 boxes, 50 steps, "
                         "craft misfit and its gradient)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint legs (checkpointed adjoints of craft misfits on "
                         "one 121x121x61 box: the field update, and the whole box-step)")
    ap.add_argument("--only-adjoint", choices=("field", "full", "nested"),
                    help="run only that adjoint leg and print its JSON (for a kernel trace "
                         "of the adjoint alone)")
    ap.add_argument("--cpu-threads", type=int,
                    default=min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 16))
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        return 2

    if a.only_adjoint:
        leg = {"field": adjoint_case, "full": full_adjoint_case,
               "nested": nested_adjoint_case}[a.only_adjoint]()
        print(json.dumps({"adjoint": {a.only_adjoint: leg}}), flush=True)
        return 0

    import torch
    from g3m3_amd import _lib
    from g3m3_amd.driver import FastRun, Run
    from g3m3_amd.workload import make_workload

    wl = make_workload(a.workload)
    run = (FastRun if a.driver == "native" else Run)(wl)
    if a.driver == "native" and a.species_kernels != "auto":
        run.set_species_kernels(a.species_kernels)
    run.init_dt()
    for _ in range(a.warmup):
        run.full_step()
    torch.cuda.synchronize()
    L = _lib.lib()
    L.g3m3_prof_select(a.prof_kernel.encode())
    cu0, bs0 = run.cell_updates, run.box_steps
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run.full_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import ctypes
    ms = ctypes.c_double(0)
    nl = ctypes.c_longlong(0)
    L.g3m3_prof_read(ctypes.byref(ms), ctypes.byref(nl))
    L.g3m3_prof_select(None)
    _lib.check("bench")
    cu = run.cell_updates - cu0
    wall = t1 - t0
    value = cu / wall
    if hasattr(run, "sync_arrays"):
        run.sync_arrays()
    for k in ("qrho", "bx", "epres"):
        v = run.a[k]
        assert bool(torch.isfinite(v).all()), f"non-finite {k} after the timed steps"

    roof = None
    traffic, tsrc = measured_traffic(a.workload, a.prof_kernel)
    if nl.value and a.prof_kernel in KERNEL_BYTES:
        per_unit, kind = KERNEL_BYTES[a.prof_kernel]
        per_launch = per_unit * units_per_launch(kind, wl)
        avg_s = ms.value / 1e3 / nl.value
        achieved = per_launch / avg_s / 1e9
        roof = {"kernel": a.prof_kernel,
                "kernels": list(STAGE_KERNELS.get(a.prof_kernel, (a.prof_kernel,))),
                "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "traffic_source": tsrc, "bytes_per_launch": per_launch,
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": nl.value}

    whole = {"bytes_per_cell_update": WHOLE_STEP_BYTES,
             "achieved": round(value * WHOLE_STEP_BYTES / 1e9, 1), "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": round(value * WHOLE_STEP_BYTES / 1e9 / HBM_PEAK_GBS, 4),
             "ms_per_box_step": round(wall / (run.box_steps - bs0) * 1e3, 4)}

    divb = None
    if not a.no_divb:
        # reference criterion: res^2 <= |b|^2 * npts * errlim^2; rel-residual 1e-6
        npts = wl.nx * wl.ny * wl.nz
        errlim = 1e-6 / np.sqrt(npts)
        info, b2 = [], []
        if hasattr(run, "sync_arrays"):
            run.sync_arrays()
        for box in range(1, wl.ngrd + 1):   # |b|^2 of each box's system, before the solves
            X, Y, Z = (run.a[k][box - 1].double() for k in ("bx", "by", "bz"))
            rx = float(wl.grid["xspac"][box - 1])
            rhs = ((X[1:-1, 1:-1, 2:] - X[1:-1, 1:-1, :-2]) + (Y[1:-1, 2:, 1:-1] - Y[1:-1, :-2, 1:-1])
                   + (Z[2:, 1:-1, 1:-1] - Z[:-2, 1:-1, 1:-1])) / (2.0 * rx)
            b2.append(float((rhs * rhs).sum()))
            del X, Y, Z, rhs
        torch.cuda.synchronize()
        d0 = time.perf_counter()
        for box in range(1, wl.ngrd + 1):
            st, it, r2 = run.divb(box, errlim, mode=a.divb_mode, max_iters=20000)
            info.append({"box": box, "status": st, "iterations": it,
                         "achieved_rel_residual": float(np.sqrt(max(r2, 0.0) / b2[box - 1]))
                         if b2[box - 1] > 0 else 0.0})
        torch.cuda.synchronize()
        dsec = time.perf_counter() - d0
        per_step = wall / a.steps
        modes = {0: "CGNR (the reference's solver), fp32 scalars",
                 1: "CGNR (the reference's solver), fp32 vectors, fp64 dots/scalars",
                 2: "CG on the same system, fp32 vectors, fp64 dots/scalars"}
        divb = {"seconds_all_boxes": round(dsec, 3), "boxes": info,
                "rel_residual_target": 1e-6,
                "mode": modes.get(a.divb_mode, str(a.divb_mode)),
                "cell_updates_per_s_incl_divb_every_10_steps":
                    cu / a.steps * 10 / (per_step * 10 + dsec)}

    prec = None if a.no_precision else precision_case(a.divb_mode)
    batch = None if a.no_batch else batch_case()
    grad = None if a.no_gradient else gradient_case()
    adj = None if a.no_adjoint else adjoint_case()
    if adj is not None:
        adj = {"field": adj, "full": full_adjoint_case(), "nested": nested_adjoint_case()}
    cpu = None if a.no_cpu else cpu_baseline(a.workload, a.cpu_threads)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dipole + solar-wind initial state)",
        "driver": a.driver,
        "config": {"workload": f"{a.workload}: {wl.ngrd} nested {wl.nx}x{wl.ny}x{wl.nz} boxes, "
                               f"spacing {list(map(float, wl.grid['xspac']))}, "
                               f"eta={float(wl.scal['eta'])}",
                   "box_steps": run.box_steps - bs0, "cell_updates": cu},
        "roofline": roof, "whole_step": whole, "cpu_baseline": cpu, "divb": divb,
    }
    if hasattr(run, "fct_shape"):
        kc, order, cands = run.fct_shape()
        out["fct_autotune"] = {"k_planes_per_block": kc, "block_order": order,
                               "candidates_us": cands}
        kind, pc = run.species_kernels()
        out["species_kernel_autotune"] = {"kernels": kind, "candidates_us": pc}
    if prec is not None:
        out["divb_precision"] = prec
    if batch is not None:
        out["batch"] = batch
    if grad is not None:
        out["gradient"] = grad
    if adj is not None:
        out["adjoint"] = adj
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
