"""The step sequence pinned to the reference's whole program.

tests/golden/main_*.json hold what the unchanged reference program
(code/3d_main.f90 with all its subroutines, built serial by oracle/Makefile)
wrote on the input files of tests/golden/main_inputs/ (five nested
121x121x61 boxes).  Here the host set-up of g3m3_amd.mainprog (the main
program's computations from its input), the oracle's initialisation block,
OracleRun (the step composition, driving the reference's compiled
subroutines) and the oracle's torus source, with the C library's powf and
expf as the reference uses them, must reproduce the fluid records bit for
bit, the box-step times of speeds.dat and the wall fluxes of fluxes.dat
exactly:

* pin_base: both full steps (62 box-steps);
* pin_tilt_torus: the first full step (31 box-steps) with the dipole
  tilting (b0 rebuilt before each of the two Lax-Wendroff steps), the torus
  injected and the ionospheric source on.

This pins the composition, the dt synchronisation and subcycling, the tilt
schedule, the torus, the set-up and the initialisation block against the
reference itself."""
import pytest

from mainpin import compare_fluxes, compare_records, compare_steps, golden, oracle_main_loop

pytestmark = pytest.mark.slow


@pytest.mark.parametrize("name,nsteps", [("pin_base", None), ("pin_tilt_torus", 1)])
def test_oracle_reproduces_the_reference_program(name, nsteps):
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    gold = golden(name)
    dims = tuple(gold["grid"])
    log = None
    seen = []
    for fname, t, recs, log in oracle_main_loop(name, nsteps):
        bad = compare_records(recs, gold["fluid"][fname], dims, exact=True)
        print(f"{name} {fname}: t = {t!r}, {len(bad)} records differ", flush=True)
        assert not bad, f"{fname}: {bad[:6]}"
        seen.append(fname)
    assert seen == sorted(gold["fluid"])[:len(seen)] and seen
    n = len(log["steps"])
    part = dict(gold, steps=gold["steps"][:n], fluxes=gold["fluxes"][:2 * len(log["fluxes"])])
    assert not compare_steps(log["steps"], part), compare_steps(log["steps"], part)
    assert not compare_fluxes(log["fluxes"], part), compare_fluxes(log["fluxes"], part)
