"""GPU parity of whole box-steps and full steps (the reference's call
sequence, code/3d_main.f90:1925-3095) against the oracle driving the same
sequence on the CPU, and between the GPU's own drivers.

Every run chooses its own time steps (no step is replayed from the other
side), and the time steps are compared after every full step, exactly.
Bar: every one of the 34 fields within rel-L2 1e-5 of the oracle after the
checked steps, and bit for bit (uint32 views), which is what the kernels
deliver.  divB cleaning (the 61x61x31 case every 10 steps) is checked on
its own at each cleaning against an all-FP64 solve of the same field, with
the bar stated there; the oracle then continues from the GPU's cleaned
field."""
import numpy as np
import pytest

from common import bit_diffs, same_bits
from g3m3_amd.workload import FIELDS, make_workload
from oracle.boxstep import OracleRun

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / max(nb, 1e-30))


def oracle_backend():
    from oracle import ref
    return "ref" if ref.available() else "restate"


def same_steps(a, b, rel=0.0):
    """t_step and t_stepnew of two runs after a full step: equal, or within rel."""
    for x, y in ((a.t_step, b.t_step), (a.t_stepnew, b.t_stepnew)):
        x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
        if rel == 0.0:
            if not np.array_equal(x, y):
                return False
        elif not np.all(np.abs(x - y) <= rel * np.abs(y)):
            return False
    return True


DIVB_BAR = 1e-5


def divb_and_check(g, o, box, tol_rel):
    """divB cleaning of one box at a step where the GPU and the oracle hold
    the same state: the GPU cleans with mode 2 (CG, FP32 vectors, FP64 dots)
    to relative residual tol_rel; the all-FP64 solve (oracle, CG to 1e-13)
    of the same field is the check.  Each corrected component must lie
    within DIVB_BAR of the largest |b| component.  The oracle then carries
    on from the GPU's cleaned field, so the steps in between stay
    comparable bit for bit.  Returns the largest relative error."""
    from oracle import restate as O
    b = box - 1
    pre = g.state_host()
    for k in ("bx", "by", "bz"):
        assert same_bits(pre[k][b], o.a[k][b]), f"box {box}: {k} differs before divB"
    stat, it, _ = g.divb(box, tol_rel, mode=2)
    assert stat == 1, f"box {box}: divB status {stat} after {it} iterations"
    ref = O.divB_correct_f64(o.a["bx"][b], o.a["by"][b], o.a["bz"][b],
                             float(o.a["xspac"][b]), method="cg")
    post = g.state_host()
    scale = max(float(np.abs(r).max()) for r in ref)
    err = max(float(np.abs(post[k][b].astype(np.float64) - r).max()) / scale
              for k, r in zip(("bx", "by", "bz"), ref))
    assert err <= DIVB_BAR, f"box {box}: divB field error {err:.3e} over {DIVB_BAR}"
    for k in ("bx", "by", "bz"):
        o.a[k][b] = post[k][b]
    return err


def run_pair(name, steps, checkpoints, native=False, divb_every=0, **over):
    from g3m3_amd.driver import FastRun, Run
    from oracle.ref import with_big_stack
    wl = make_workload(name, **over)
    g = (FastRun if native else Run)(wl)
    g.init_dt()
    be = oracle_backend()
    o = OracleRun(wl, be)
    with_big_stack(o.init_dt)
    assert same_steps(g, o), "initial time steps differ"
    report = {}
    for st in range(1, steps + 1):
        with_big_stack(o.full_step)
        g.full_step()
        assert same_steps(g, o), f"step {st}: time steps {g.t_step} / {o.t_step}"
        if divb_every and st % divb_every == 0:
            for box in range(wl.ngrd, 0, -1):
                err = divb_and_check(g, o, box, wl.divb_tol)
                print(f"step {st} box {box}: divB field error {err:.2e}", flush=True)
        if st in checkpoints:
            gh = g.state_host()
            errs = {k: rel_l2(gh[k], o.a[k]) for k in FIELDS}
            report[st] = errs
            nbits = {k: bit_diffs(gh[k], o.a[k]) for k in FIELDS}
            print(f"{name} step {st} ({be}, {'native' if native else 'per-subroutine'}): "
                  f"worst field {max(errs, key=errs.get)} rel-L2 {max(errs.values()):.3e}, "
                  f"{sum(nbits.values())} values differ in their bits", flush=True)
            bad = {k: v for k, v in errs.items() if not v <= TOL}
            assert not bad, f"step {st}: fields over {TOL}: {bad}"
            bad = {k: n for k, n in nbits.items() if n}
            assert not bad, f"step {st}: fields not bit-equal: {bad}"
            if native:
                # the runtime rotates state/old; after sync both arrays and
                # wrk must hold what the reference's arrays hold
                for pre in ("old", "wrk"):
                    gp = g.state_host(pre)
                    e2 = {k: rel_l2(gp[k], o.a[pre + k]) for k in FIELDS}
                    bad = {k: v for k, v in e2.items() if not v <= TOL}
                    assert not bad, f"step {st}: {pre} arrays over {TOL}: {bad}"
    assert g.box_steps == o.box_steps
    return report


def test_tiny_nested_bitexact():
    """Two nested boxes, 3 full steps: flanks, corer and subcycling."""
    run_pair("tiny", 3, {1, 2, 3})


def test_small_nested_eta():
    """Two nested 21x21x11 boxes with the uniform-resistivity extension."""
    run_pair("small", 4, {1, 4}, eta=1e-3)


def test_single61_1_10():
    """The single 61x61x31 box (the reference's CPU-runnable case)."""
    run_pair("single61", 10, {1, 10})


@pytest.mark.parametrize("name,steps,over", [("tiny", 3, {}), ("small", 4, {"eta": 1e-3}),
                                             ("single61", 10, {})])
def test_native_runtime(name, steps, over):
    """The native runtime (fused kernels, one call per full step) against
    the oracle: same bar."""
    run_pair(name, steps, {1, steps}, native=True, **over)


def test_native_equals_per_subroutine_three_boxes():
    """Three nested 25x25x13 boxes, 3 full steps: the native runtime (fused
    kernels, buffer rotation, flanks reading the coarse values from a
    gathered slab) and the per-subroutine path (one call per reference
    subroutine, coarse values read in place) agree bit for bit in every
    state, old and wrk array after each step."""
    from g3m3_amd.driver import FastRun, Run
    wl = make_workload("small", ngrd=3, steps=3)
    a, b = FastRun(wl), Run(wl)
    a.init_dt()
    b.init_dt()
    for st in range(1, 4):
        print(f"three boxes, step {st}", flush=True)
        b.full_step()
        a.full_step()
        assert same_steps(a, b), f"step {st}: time steps differ"
        for pre in ("", "old", "wrk"):
            ha, hb = a.state_host(pre), b.state_host(pre)
            bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
            assert not bad, f"step {st} {pre or 'state'}: {bad}"
    assert a.box_steps == b.box_steps


def test_native_equals_per_subroutine_81x81x41_three_boxes():
    """The forward run of the differentiation case's geometry (three nested
    81x81x41 boxes, resistivity 1e-3): native runtime and per-subroutine
    path bit-equal in every state, old and wrk array after each of 2 full
    steps (14 box-steps)."""
    from g3m3_amd.driver import FastRun, Run
    wl = make_workload("nested121", L=40, ngrd=3, eta=1e-3)
    assert (wl.nx, wl.ny, wl.nz, wl.ngrd) == (81, 81, 41, 3)
    a, b = FastRun(wl), Run(wl)
    a.init_dt()
    b.init_dt()
    for st in range(1, 3):
        print(f"81x81x41 x3, step {st}", flush=True)
        b.full_step()
        a.full_step()
        assert same_steps(a, b), f"step {st}: time steps differ"
        for pre in ("", "old", "wrk"):
            ha, hb = a.state_host(pre), b.state_host(pre)
            bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
            assert not bad, f"step {st} {pre or 'state'}: {bad}"
    assert a.box_steps == b.box_steps


def test_native_equals_per_subroutine_nested121_50_steps():
    """The benchmark configuration itself -- five nested 121x121x61 boxes
    with resistivity, 50 full steps -- through the native runtime (fused
    kernels, autotuned launch shapes, buffer rotation, gathered flank slabs)
    and through the per-subroutine path: every state field bit-equal after
    steps 1, 10 and 50, and old and wrk after step 50."""
    from g3m3_amd.driver import FastRun, Run
    wl = make_workload("nested121")
    a, b = FastRun(wl), Run(wl)
    a.init_dt()
    b.init_dt()
    for st in range(1, 51):
        b.full_step()
        a.full_step()
        assert same_steps(a, b), f"step {st}: time steps differ"
        if st in (1, 10, 50):
            print(f"nested121 step {st}", flush=True)
            for pre in (("", "old", "wrk") if st == 50 else ("",)):
                ha, hb = a.state_host(pre), b.state_host(pre)
                bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
                assert not bad, f"step {st} {pre or 'state'}: {bad}"
    assert a.box_steps == b.box_steps


def test_native_single61_100_steps():
    """The 61x61x31 single-grid case over 100 steps, every field checked
    after 1, 10 and 100 steps (Lapidus and FCT every step), bit for bit."""
    run_pair("single61", 100, {1, 10, 100}, native=True)


def test_native_single61_100_steps_divb():
    """The 61x61x31 configuration as specified: 100 steps with Lapidus and
    FCT every step and divB cleaning every 10 steps (mode 2, relative
    residual 1e-6).  Each cleaning is checked against the all-FP64 solve of
    the same field (every component within 1e-5 of the largest |b|), the
    oracle continues from the GPU's cleaned field, and every step is
    bit-equal with equal time steps; fields checked after steps 10, 50,
    100."""
    run_pair("single61", 100, {10, 50, 100}, native=True, divb_every=10)


def test_fluid_restart_round_trip(tmp_path):
    """A run written to a fluid file mid-run and read into a fresh run of
    the same shape holds the same state, time and planet lists; both then
    advance identically (a restart in the reference zeroes old and wrk,
    and so does read_fluid)."""
    from g3m3_amd.driver import FastRun
    wl = make_workload("tiny")
    a = FastRun(wl)
    a.init_dt()
    a.full_step()
    path = str(tmp_path / "fluid11")
    a.write_fluid(path)
    b = FastRun(make_workload("tiny", seed=99))   # different initial state
    b.read_fluid(path)
    assert b.t == a.t
    ha, hb = a.state_host(), b.state_host()
    assert all(same_bits(ha[k], hb[k]) for k in FIELDS)
    a.read_fluid(path)      # the same restart semantics on both sides
    for _ in range(2):
        a.full_step()
        b.full_step()
    ha, hb = a.state_host(), b.state_host()
    bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
    assert not bad, bad


def test_fct_launch_shapes_equal():
    """Every FCT launch shape the autotuner chooses among (8, 16 or 32
    k-planes per block; XCD-contiguous or plain block order) gives the same
    bits over whole steps of a nested run; the tuner's pick is one of them."""
    from g3m3_amd.driver import FastRun
    ref = None
    for kc in (8, 16, 32):
        for order in ("xcd", "plain"):
            r = FastRun(make_workload("small", eta=1e-3))
            if ref is None:
                kc0, order0, cands = r.fct_shape()
                assert kc0 in (8, 16, 32) and order0 in ("xcd", "plain") and len(cands) == 6
            r.set_fct_shape(kc, order)
            r.init_dt()
            print(f"FCT shape {kc}/{order}", flush=True)
            for _ in range(2):
                r.full_step()
            h = r.state_host()
            if ref is None:
                ref = h
            else:
                bad = [k for k in FIELDS if not same_bits(h[k], ref[k])]
                assert not bad, f"shape {kc}/{order} differs on {bad}"


def test_species_kernels_equal():
    """The two families of species push and Lapidus kernels the autotuner
    chooses between (k-marching LDS tiles, one thread per cell) give the same
    bits over whole steps of a nested run with the resistivity extension;
    the tuner keeps the faster."""
    from g3m3_amd.driver import FastRun
    ref = None
    for kind in ("cell", "march"):
        r = FastRun(make_workload("small", ngrd=3, eta=1e-3))
        k0, cands = r.species_kernels()
        assert k0 in ("cell", "march") and set(cands) == {"cell", "march"}
        assert cands[k0] == min(cands.values()), (k0, cands)   # the faster one
        r.set_species_kernels(kind)
        r.init_dt()
        print(f"species kernels {kind}", flush=True)
        for _ in range(3):
            r.full_step()
        h = {p: r.state_host(p) for p in ("", "old", "wrk")}
        if ref is None:
            ref = h
        else:
            for p in h:
                bad = [k for k in FIELDS if not same_bits(h[p][k], ref[p][k])]
                assert not bad, f"{p or 'state'} differs on {bad}"


@pytest.mark.parametrize("dims,spacing", [((64, 66, 18), 1.0), ((63, 65, 17), 0.75)])
def test_species_kernels_equal_tile_edges(dims, spacing):
    """Single boxes whose interior is an exact multiple of the k-marching
    tiles (62 x 4 cells, 16 planes: the last tile's halo lane owns the x face)
    and one cell short of it, with a grid step that is not a power of two
    (true divisions instead of reciprocal products): both species-kernel
    families give the same bits in every state, old and wrk array."""
    from g3m3_amd.driver import FastRun
    ref = None
    for kind in ("cell", "march"):
        r = FastRun(make_workload("single61", dims=dims, spacing=spacing))
        r.set_species_kernels(kind)
        r.init_dt()
        print(f"{dims} spacing {spacing}: {kind}", flush=True)
        for _ in range(3):
            r.full_step()
        h = {p: r.state_host(p) for p in ("", "old", "wrk")}
        if ref is None:
            ref = h
        else:
            for p in h:
                bad = [k for k in FIELDS if not same_bits(h[p][k], ref[p][k])]
                assert not bad, f"{p or 'state'} differs on {bad}"


TILT = dict(tilting=True, tilt=12.0, dtilt=2.0, ut=3.0, t_equiv=2.5, planet_per=10.0)


def test_tilting_native_equals_per_subroutine():
    """Dipole tilting during the run (code/3d_main.f90:1966-1977,
    2173-2187, 2465-2478) on three nested boxes: every box-step recomputes
    its b0 for the tilt and rotation at its half step before step 1 and at
    its full step before step 2.  Native runtime and per-
    subroutine path bit-equal in b0 and every state, old and wrk array after
    each of 3 steps, and b0 does change."""
    from g3m3_amd.driver import FastRun, Run
    wl = make_workload("small", ngrd=3, **TILT)
    a, b = FastRun(wl), Run(wl)
    b0_start = b.a["bx0"].cpu().numpy().copy()
    a.init_dt()
    b.init_dt()
    for st in range(1, 4):
        print(f"tilting, step {st}", flush=True)
        b.full_step()
        a.full_step()
        assert same_steps(a, b), f"step {st}: time steps differ"
        for pre in ("", "old", "wrk"):
            ha, hb = a.state_host(pre), b.state_host(pre)
            bad = [k for k in FIELDS if not same_bits(ha[k], hb[k])]
            assert not bad, f"step {st} {pre or 'state'}: {bad}"
        for k in ("bx0", "by0", "bz0"):
            assert same_bits(a.a[k].cpu().numpy(), b.a[k].cpu().numpy()), (st, k)
    assert not np.array_equal(b.a["bx0"].cpu().numpy(), b0_start)


def test_tilting_against_oracle():
    """The per-subroutine path with tilting against the oracle driving the
    same sequence (the reference build's mak_dip_grd where it is built):
    b0 and every field bit-equal after each of 3 steps."""
    from g3m3_amd.driver import Run
    from oracle.ref import with_big_stack
    wl = make_workload("tiny", **TILT)
    g = Run(wl)
    g.init_dt()
    o = OracleRun(wl, oracle_backend())
    with_big_stack(o.init_dt)
    for st in range(1, 4):
        with_big_stack(o.full_step)
        g.full_step()
        print(f"tilting vs oracle, step {st}", flush=True)
        assert same_steps(g, o), f"step {st}: time steps differ"
        for k in ("bx0", "by0", "bz0"):
            assert same_bits(g.a[k].cpu().numpy(), o.a[k]), (st, k)
        gh = g.state_host()
        errs = {k: rel_l2(gh[k], o.a[k]) for k in FIELDS}
        assert max(errs.values()) <= TOL, errs
        bad = [k for k in FIELDS if not same_bits(gh[k], o.a[k])]
        assert not bad, f"step {st}: not bit-equal: {bad}"


@pytest.mark.parametrize("native", [True, False])
def test_collapsed_step_stops_the_run(native):
    """A time step below t_step_min = 1e-6, or one that is not finite, stops
    the run with an error before any box-step, as the main loop stops
    (code/3d_main.f90:1969-1973); the state is unchanged."""
    from g3m3_amd import _lib
    from g3m3_amd.driver import FastRun, Run
    wl = make_workload("tiny")
    for dt in (1e-7, float("nan")):
        g = (FastRun if native else Run)(wl)
        g.init_dt()
        before = g.state_host()
        print(f"collapsed step dt={dt} native={native}", flush=True)
        with pytest.raises(_lib.G3M3Error):
            g.full_step(dt_override=[dt] * wl.ngrd)
        after = Run.state_host(g)     # the arrays as they stand (no sync on a failed run)
        assert g.box_steps == 0
        assert all(np.array_equal(before[k], after[k]) for k in FIELDS)
