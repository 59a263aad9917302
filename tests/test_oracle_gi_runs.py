"""The oracle's global-implicit time loop (oracle.gi_column_run: the oracle's
residual and Jacobian, direct Newton solves, PETSc's SNES tests, the
reference's step controller in pflotran_amd.stepper) against the
reference's run golds: the 1-cell batch kinetics decks
(ascem/batch/calcite-kinetics{,-volume-fractions}) and the 1-D columns
(ascem/1d/1d-calcite, shortcourse/1D_Calcite/calcite_tran_only,
default/column/valocchi_ionx) and the linear K_D isotherm decks
(default/column/solute_KD, default/batch/solute_KD_w{,o}_mineral).  Values
to the suites' 1e-12 absolute, the solution's 2-norm to the gold's digits,
and the step and Newton-iteration counts exactly: each step's size follows the
previous step's Newton count, so the counts pin the SNES tests, the growth
rule, the waypoints (PERIODIC TIME 1 y in 1d-calcite, TIMES in
calcite_tran_only) and the final-time landing."""
import pytest

import oracle as orc
from golden_runs import RUN_DECKS, differences, load, outputs, solution_norm, within

TOL = 1e-12


@pytest.mark.parametrize("name", sorted(RUN_DECKS))
def test_oracle_run_reproduces_the_gold(name):
    deck, db, g = load(name)
    ts, cells, t = orc.gi_column_run(deck, db)
    sol = g["SOLUTION: Transport"]
    assert ts.steps == sol["Time Steps"]
    assert sum(h[1] for h in ts.history) == sol["Newton Iterations"]
    assert sum(h[2] for h in ts.history) == sol["Time Step Cuts"]
    diffs = differences(g, outputs(t, cells.arrays, deck.water_density(), deck.chemistry.molal))
    bad = [d for d in diffs if not within(name, d[1], d[3], d[4])]
    assert len(diffs) >= 1 and not bad, bad[:3]
    # the 2-norm of the final solution vector (the free-ion molalities of
    # every cell), to the 14 digits the gold prints
    assert abs(solution_norm(cells.arrays) / sol["Solution 2-Norm"] - 1.0) <= 1e-13


def test_wrong_growth_rule_misses_the_schedule():
    # a plausible mistake planted in the oracle's loop: growth from a Newton
    # count one too low leaves the vf deck's first step un-halved and the
    # run takes 61 steps, not the gold's 62
    from pflotran_amd import stepper
    deck, db, g = load("calcite-kinetics-volume-fractions")
    orig = stepper.TimeStepper.grow
    try:
        stepper.TimeStepper.grow = lambda self, dt, n: orig(self, dt, n - 5)
        ts, _, _ = orc.gi_column_run(deck, db)
    finally:
        stepper.TimeStepper.grow = orig
    assert ts.steps != g["SOLUTION: Transport"]["Time Steps"]
