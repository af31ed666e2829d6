"""The whole-box-step adjoint through nested boxes.

A stack of boxes, innermost first (the reference's box 1 .. n_grids), each
the device model of g3m3_amd.adjoint.FullModel: the outermost box is bounded
by bndry_outer and carries the controls' inflow; every inner box is bounded
by the flanks of its parent (code/sub_flanks.f90: trilinear in space, linear
in time between the parent's states at the start and the end of its last
step) and, whenever its time catches up with its parent's, restricted into
the parent's core (code/sub_corer.f90).  The boxes step on the reference's
subcycling schedule (code/3d_main.f90:1925-3095) with fixed time steps, each
box's twice its inner neighbour's: per full step the boxes are visited from
the outermost in, 2^(n-1) times, and a box steps when it has caught up with
the innermost one (``nested_schedule``).

``NestedRollout`` is a ``torch.autograd.Function`` from (ctl, the boxes'
states at step 0) to the craft samples every ``every`` full steps (findgrid
picks the finest box holding each craft, code/cft_find_aux_cube.f90).  The
forward pass keeps the boxes' states every ``ckpt`` full steps, and when a
gradient is wanted the last segment's state versions and tapes; the
backward pass recomputes every earlier segment from its checkpoint, keeping
every state version and a tape per box-step, and runs the hand-written
adjoints back through the schedule: a box-step's adjoint (``g3m3_full_model_adjoint`` /
``g3m3_full_model_adjoint_inner``, which also adds the flanks' adjoints into
the parent's two states), the corer's (``g3m3_full_model_corer_adjoint``)
and the sampling's (``g3m3_sample_fields_adjoint``).
"""
from __future__ import annotations

import ctypes as _ct

import numpy as np

from . import _lib
from .adjoint import (FULL_CTL_NAMES, FULL_SAMPLED, NSTATE, FullModel, FullSetup, _modes,
                      dipole_b0, full_setup)

F = np.float32


def _trilinear(src: np.ndarray, x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """src (..., nz, ny, nx) at fractional node coordinates x (nx'), y (ny'),
    z (nz') inside it: (..., nz', ny', nx'), float64."""
    def axis(c, n):
        i = np.clip(np.floor(c).astype(np.int64), 0, n - 2)
        return i, c - i
    nz, ny, nx = src.shape[-3:]
    i, fx = axis(x, nx)
    j, fy = axis(y, ny)
    k, fz = axis(z, nz)
    s = src.astype(np.float64)
    out = 0.0
    for dk, wz in ((0, 1.0 - fz), (1, fz)):
        for dj, wy in ((0, 1.0 - fy), (1, fy)):
            for di, wx in ((0, 1.0 - fx), (1, fx)):
                v = s[..., (k + dk)[:, None, None], (j + dj)[None, :, None], (i + di)[None, None, :]]
                out = out + v * (wz[:, None, None] * wy[None, :, None] * wx[None, None, :])
    return out


def nested_setup(nbox: int = 2, nx: int = 20, ny: int = 16, nz: int = 12,
                 spacing: float = 1.0, seed: int = 20261209, **kw) -> list[FullSetup]:
    """``nbox`` boxes of nx x ny x nz cells, innermost first: the outermost
    is full_setup(nx, ny, nz, spacing, seed, **kw); each inner box has half
    its parent's spacing and time step, its lower corner on a parent node
    about a quarter of the parent in from each lower face (centred), its
    state the trilinear interpolation of the parent's with +-2% smooth
    variations, and b0 the dipole at its own nodes."""
    out = [full_setup(nx, ny, nz, spacing, seed, **kw)]
    for _ in range(nbox - 1):
        P = out[-1]
        H = F(P.scal["rx"])
        h = F(H / F(2.0))
        off = [int(round((n - 1) / 4.0)) for n in (nx, ny, nz)]
        lo = [F(F(P.scal[a + "min"]) + F(o) * H) for a, o in zip("xyz", off)]
        ax = (lo[0] + h * np.arange(nx, dtype=F))[None, None, :]
        ay = (lo[1] + h * np.arange(ny, dtype=F))[None, :, None]
        az = (lo[2] + h * np.arange(nz, dtype=F))[:, None, None]
        fx = off[0] + 0.5 * np.arange(nx)
        fy = off[1] + 0.5 * np.arange(ny)
        fz = off[2] + 0.5 * np.arange(nz)
        st = _trilinear(P.state, fx, fy, fz)
        # +-2% smooth variations per array on the finer grid: the
        # interpolated fields alone are piecewise linear, whose exactly flat
        # second differences would leave flux-correction limiter branches to
        # rounding
        rng = np.random.default_rng(seed + 7 * len(out))
        for a in range(st.shape[0]):
            st[a] *= 1.0 + 0.02 * _modes(rng, st.shape[1:], float(h))
        st = st.astype(F)
        scal = dict(P.scal)
        hi = [F(l + h * F(n - 1)) for l, n in zip(lo, (nx, ny, nz))]
        scal.update(rx=h, delt=F(F(P.scal["delt"]) / F(2.0)), xmin=lo[0], xmax=hi[0],
                    ymin=lo[1], ymax=hi[1], zmin=lo[2], zmax=hi[2])
        grid = dict(grid_minvals=np.array([lo], F), grid_maxvals=np.array([hi], F),
                    xspac=np.array([h], F), re_equiv=P.grid["re_equiv"])
        out.append(FullSetup(nx=nx, ny=ny, nz=nz, state=np.ascontiguousarray(st),
                             b0=dipole_b0(ax, ay, az), ctl=P.ctl.copy(), scal=scal, grid=grid))
    return out[::-1]


def nested_schedule(ts) -> list:
    """The box-steps and corer calls of one full step (code/3d_main.f90:
    1925-3095 with the time steps fixed): ts[b] is box b+1's time step
    (float32 values, each twice the one before).  Events: ("step", b, t12)
    -- box b+1 steps; t12 = (t1, t2) of the flanks at step 1's t_grid and at
    the later stages' (:2320-2330, :2589-2599; code/sub_flanks.f90:21-22),
    None for the outermost box -- and ("corer", b) -- box b+1's state into
    its parent's core (:3055-3063 after a box catches up with its parent,
    :3080-3090 for every inner box at the end of the full step)."""
    n = len(ts)
    ts = [float(F(t)) for t in ts]
    for b in range(1, n):
        if ts[b] != 2.0 * ts[b - 1]:
            raise ValueError("each box's time step must be twice its inner neighbour's")
    t_old, t_new = [0.0] * n, [0.0] * n
    ev = []
    for _ in range(int(ts[n - 1] / ts[0])):
        for b in range(n - 1, -1, -1):
            if b > 0 and not (t_old[b] == t_new[b] and t_new[b] < ts[n - 1]
                              and abs(t_new[b] - t_new[0]) <= 0.005 * ts[0]):
                continue
            t_old[b] = t_new[b]
            t_new[b] = t_old[b] + ts[b]
            t12 = None
            if b < n - 1:
                p = b + 1
                delt = F(ts[b])
                w = []
                for dt in (float(delt / F(2.0)), float(delt)):
                    tg = t_old[b] + dt
                    tg = t_new[p] if tg > t_new[p] else tg
                    tg = t_old[p] if tg < t_old[p] else tg
                    t1 = F((t_new[p] - tg) / (t_new[p] - t_old[p]))
                    w += [t1, F(1.0) - t1]
                t12 = tuple(w)
            ev.append(("step", b, t12))
            if b < n - 1 and abs(t_new[b] - t_new[b + 1]) <= 0.005 * ts[b]:
                ev.append(("corer", b))
                t_old[b + 1] = t_new[b + 1]
    ev += [("corer", b) for b in range(n - 1)]
    return ev


class NestedFullModel:
    """The device models of a nested_setup stack (innermost first)."""

    def __init__(self, setups: list[FullSetup]):
        self.setups = list(setups)
        self.n = len(self.setups)
        nx, ny, nz = setups[0].nx, setups[0].ny, setups[0].nz
        if any((s.nx, s.ny, s.nz) != (nx, ny, nz) for s in setups):
            raise ValueError("every box of a nested stack has the same extents in cells")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.V = nx * ny * nz
        self.boxes = [FullModel(s) for s in self.setups]
        for b in range(self.n - 1):
            self.boxes[b].set_parent(self.setups[b + 1])
        self.events = nested_schedule([s.scal["delt"] for s in self.setups])
        self.box_steps = sum(1 for e in self.events if e[0] == "step")
        self._grid = (np.ascontiguousarray(np.concatenate([s.grid["grid_minvals"]
                                                           for s in setups])),
                      np.ascontiguousarray(np.concatenate([s.grid["grid_maxvals"]
                                                           for s in setups])),
                      np.ascontiguousarray(np.concatenate([s.grid["xspac"] for s in setups])),
                      F(setups[0].grid["re_equiv"]))

    def close(self):
        for m in self.boxes:
            m.close()

    def new_states(self):
        return [m.new_state() for m in self.boxes]

    def full_step(self, states, ctl, rec=None):
        """One full step from ``states`` (innermost first, (35, V) each;
        not changed); returns the new states.  With ``rec`` a list, every
        state version is kept and each box-step writes a tape: rec receives
        ("step", b, s_in, s_out, p_old, p_cur, t12, tape) and ("corer", b,
        child, parent_in, parent_out) for adjoint_full_step."""
        import torch
        cur, old = list(states), [None] * self.n
        for ev in self.events:
            if ev[0] == "step":
                _, b, t12 = ev
                m = self.boxes[b]
                out = torch.empty_like(cur[b])
                tape = m.new_tape() if rec is not None else None
                if t12 is None:
                    m.step(cur[b], out, ctl, tape)
                    po = pc = None
                else:
                    po, pc = old[b + 1], cur[b + 1]
                    m.step_inner(cur[b], out, ctl, po, pc, t12, tape)
                if rec is not None:
                    rec.append(("step", b, cur[b], out, po, pc, t12, tape))
                old[b], cur[b] = cur[b], out
            else:
                b = ev[1]
                # every box steps before any corer writes its core, so the
                # parent's state here is one this full step made
                parent = cur[b + 1]
                if rec is not None:
                    parent = parent.clone()
                    rec.append(("corer", b, cur[b], cur[b + 1], parent))
                    cur[b + 1] = parent
                self.boxes[b].corer(cur[b], parent)
        return cur

    def adjoint_full_step(self, rec, finals, bars, ctl_bar, ctl):
        """The adjoint of the full step recorded in ``rec`` whose outputs
        were ``finals``: from their adjoints ``bars`` (per box, consumed)
        the adjoints of its inputs (per box); ctl_bar (9 device doubles) +=
        the controls' adjoints."""
        import torch
        acc = {id(s): g for s, g in zip(finals, bars)}

        def bar(t):
            if id(t) not in acc:
                acc[id(t)] = torch.zeros_like(t)
            return acc[id(t)]

        for r in reversed(rec):
            if r[0] == "step":
                _, b, s_in, s_out, po, pc, t12, tape = r
                m = self.boxes[b]
                ob = bar(s_out)
                ib = torch.empty_like(s_in)
                if t12 is None:
                    m.adjoint(tape, ob, ib, ctl_bar, ctl)
                else:
                    m.adjoint_inner(tape, ob, ib, ctl_bar, ctl, t12, bar(po), bar(pc))
                del acc[id(s_out)]
                if id(s_in) in acc:
                    acc[id(s_in)] += ib
                else:
                    acc[id(s_in)] = ib
            else:
                _, b, child, p_in, p_out = r
                pb = bar(p_out)
                self.boxes[b].corer_adjoint(bar(child), pb)
                del acc[id(p_out)]
                # the corer's input state is read by nothing after it
                assert id(p_in) not in acc
                acc[id(p_in)] = pb
        first = {}
        for r in rec:            # each box's input version: its first step's s_in
            if r[0] == "step" and r[1] not in first:
                first[r[1]] = r[2]
        return [bar(first[b]) for b in range(self.n)]

    # --- craft samples over the stack (findgrid: the finest box holding the point) ---
    def _stack(self, states):
        import torch
        return torch.stack([st[list(FULL_SAMPLED)] for st in states], dim=1).contiguous()

    def sample(self, states, pos, out, gp):
        buf = self._stack(states)
        ptrs = (_ct.c_void_p * len(FULL_SAMPLED))(*[buf[q].data_ptr()
                                                    for q in range(len(FULL_SAMPLED))])
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields", pos.shape[0], pos, self.n, mn, mx, xs, re, self.nx, self.ny,
                  self.nz, len(FULL_SAMPLED), ptrs, out, gp)

    def sample_adjoint(self, pos, s_bar, st_bars):
        import torch
        buf = torch.zeros((len(FULL_SAMPLED), self.n, self.V), dtype=torch.float32,
                          device=st_bars[0].device)
        ptrs = (_ct.c_void_p * len(FULL_SAMPLED))(*[buf[q].data_ptr()
                                                    for q in range(len(FULL_SAMPLED))])
        mn, mx, xs, re = self._grid
        _lib.call("sample_fields_adjoint", pos.shape[0], pos, self.n, mn, mx, xs, re, self.nx,
                  self.ny, self.nz, len(FULL_SAMPLED), ptrs, s_bar)
        idx = torch.tensor(FULL_SAMPLED, device=buf.device)
        for b in range(self.n):
            st_bars[b].index_add_(0, idx, buf[:, b])


def _nested_function():
    import torch

    class NestedRollout(torch.autograd.Function):
        """(ctl, box states at step 0...) -> the FULL_SAMPLED quantities at
        the craft after every ``every``-th of ``steps`` full steps, shape
        (steps // every, npts, 8).  ctl: float64 CPU tensor of 9 (rounded to
        float32 for the device); states: (35, V) float32 device tensors,
        innermost box first.  Backward: the gradients in ctl (accumulated in
        double on the device) and in each box's state."""

        @staticmethod
        def forward(ctx, ctl, model, steps, every, pos, ckpt, keep, *states):
            c = ctl.detach().cpu().numpy()
            ns = steps // every
            out = torch.empty((ns, pos.shape[0], len(FULL_SAMPLED)), dtype=torch.float32,
                              device="cuda")
            gp = torch.empty((pos.shape[0], 4), dtype=torch.int32, device="cuda")
            cur = [s.detach() for s in states]
            ckpts = []
            last = (steps - 1) // ckpt * ckpt
            recs = [] if keep else None
            for t in range(steps):
                if t % ckpt == 0:
                    ckpts.append([s.clone() for s in cur])
                rec = [] if keep and t >= last else None
                nxt = model.full_step(cur, c, rec)
                if rec is not None:
                    recs.append((rec, nxt))
                cur = nxt
                if (t + 1) % every == 0:
                    model.sample(cur, pos, out[(t + 1) // every - 1], gp)
            ctx.model, ctx.steps, ctx.every, ctx.ckpt = model, steps, every, ckpt
            ctx.c, ctx.pos, ctx.ckpts, ctx.last_recs = c, pos, ckpts, recs
            ctx.ctl_meta = (ctl.dtype, ctl.device, ctl.shape)
            return out

        @staticmethod
        def backward(ctx, gs):
            model, steps, every, ckpt = ctx.model, ctx.steps, ctx.every, ctx.ckpt
            gs = gs.contiguous()
            bars = [torch.zeros((NSTATE, model.V), dtype=torch.float32, device="cuda")
                    for _ in range(model.n)]
            ctl_bar = torch.zeros(len(FULL_CTL_NAMES), dtype=torch.float64, device="cuda")
            for seg in reversed(range(0, steps, ckpt)):
                end = min(seg + ckpt, steps)
                recs = ctx.last_recs
                ctx.last_recs = None
                if recs is None:
                    # recompute the segment from its checkpoint (the last
                    # segment's records come from the forward pass)
                    cur = ctx.ckpts[seg // ckpt]
                    recs = []
                    for t in range(seg, end):
                        rec = []
                        nxt = model.full_step(cur, ctx.c, rec)
                        recs.append((rec, nxt))
                        cur = nxt
                for t in reversed(range(seg, end)):
                    rec, finals = recs.pop()
                    if (t + 1) % every == 0:
                        model.sample_adjoint(ctx.pos, gs[(t + 1) // every - 1], bars)
                    bars = model.adjoint_full_step(rec, finals, bars, ctl_bar, ctx.c)
                    del rec, finals
            dtype, device, shape = ctx.ctl_meta
            g_ctl = ctl_bar.cpu().to(dtype).reshape(shape).to(device)
            return (g_ctl, None, None, None, None, None, None, *bars)

    return NestedRollout


_NESTED = None


def nested_rollout(ctl, states, model: NestedFullModel, steps: int, every: int, pos,
                   ckpt: int = 5):
    """Craft samples of the nested stack's rollout, differentiable in ctl
    and every box's initial state."""
    import torch
    global _NESTED
    if _NESTED is None:
        _NESTED = _nested_function()
    keep = torch.is_grad_enabled() and (ctl.requires_grad or any(s.requires_grad for s in states))
    return _NESTED.apply(ctl, model, steps, every, pos, ckpt, keep, *states)
