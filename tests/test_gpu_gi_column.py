"""The reference's run golds through the device time loop
(pflotran_amd.gi_run.ColumnRun: the reference's step controller around
pfx_gi_solver_timestep): the 1-cell batch kinetics decks
(ascem/batch/calcite-kinetics{,-volume-fractions}), the 20-cell 1-D
calcite column (ascem/1d/1d-calcite), the 100-cell shortcourse calcite
column (shortcourse/1D_Calcite/calcite_tran_only), the 100-cell Valocchi
ion-exchange column (default/column/valocchi_ionx) and the linear K_D
isotherm decks (default/column/solute_KD, 100 cells, and the 1-cell
default/batch/solute_KD_w{,o}_mineral) and the mineral-prefactor column
(default/column/mineral_prefactor: rate prefactors, a uranium basis).
Values at the gold's cells to the regression suites' criteria (1e-12
absolute; the prefactor column's relaxed ones), the solution's 2-norm to the gold's
digits; the step count and the Newton iteration count of the whole run
equal the gold's (they fix the schedule: each step's size follows the
previous step's Newton count)."""
import pytest

from golden_runs import RUN_DECKS, differences, load, outputs, solution_norm, within
from pflotran_amd import gi_run

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("name", sorted(RUN_DECKS))
def test_run_gold_through_the_device_time_loop(name):
    deck, db, g = load(name)
    run = gi_run.ColumnRun(deck, db)
    try:
        res = run.run()
    finally:
        run.close()
    sol = g["SOLUTION: Transport"]
    print(f"{name}: {res.steps} steps ({sol['Time Steps']:.0f} in the gold), {res.newton_iterations} Newton "
          f"iterations ({sol['Newton Iterations']:.0f}), {res.cuts} cuts")
    diffs = differences(g, outputs(run.t, res.state, run.den, deck.chemistry.molal))
    for d in diffs[:6]:
        print("  %.3e  %s cell %s: %.14e gold %.14e" % d)
    bad = [d for d in diffs if not within(name, d[1], d[3], d[4])]
    assert len(diffs) >= 1 and not bad, bad[:3]
    assert abs(solution_norm(res.state) / sol["Solution 2-Norm"] - 1.0) <= 1e-13
    assert res.steps == sol["Time Steps"] and res.cuts == sol["Time Step Cuts"]
    assert res.newton_iterations == sol["Newton Iterations"]
