"""The oracle against the reference's own regression golds
(regression_tests/ascem/batch/*.regression.gold, 1e-12 absolute per batch.cfg).
A one-cell global-implicit run without transport solves exactly the per-cell
reaction system of RReact, so these pin the oracle's chemistry end to end."""
import numpy as np
import pytest

import oracle as orc
from chemdefs import calcite_kinetics, ca_carbonate, carbonate, valocchi, valocchi_sorbed
from conftest import gold
from pflotran_amd import abi
from pflotran_amd.database import water_density_ifc67

TOL = 1e-12  # batch.cfg default-test-criteria
DEN = water_density_ifc67()


def speciate(t, setup):
    cells = orc.HostCells.allocate(t, 1)
    if "vf" in setup:
        cells["mnrl_volfrac"][0, 0] = setup["vf"]
        cells["mnrl_area"][0, 0] = setup["area"]
    o = abi.default_opts(porosity=setup["porosity"], den_kg=DEN)
    rc, its = orc.equilibrate(t, setup["ctype"], np.array(setup["conc"]), cells, o, rxn=setup.get("rxn"))
    assert rc == 0 and its[0] > 0
    return cells, o


@pytest.mark.parametrize("act", [False, True])
def test_carbonate_speciation(golds, act):
    name = "carbonate-debye-huckel-activity" if act else "carbonate-unit-activity"
    t, setup = carbonate(act)
    cells, _ = speciate(t, setup)
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    assert abs(-np.log10(g[0] * m[0]) - gold(golds, name, "GENERIC: pH")) < TOL
    assert abs(cells["total"][0][0] - gold(golds, name, "CONCENTRATION: Total H+")) < TOL
    assert abs(cells["total"][0][1] - gold(golds, name, "CONCENTRATION: Total HCO3-")) < TOL
    if act:
        assert abs(g[0] - gold(golds, name, "GENERIC: Gamma H+")) < TOL
        assert abs(g[1] - gold(golds, name, "GENERIC: Gamma HCO3-")) < TOL
    else:
        assert np.all(g == 1.0)


@pytest.mark.parametrize("act", [False, True])
def test_ca_carbonate_speciation(golds, act):
    name = "ca-carbonate-debye-huckel-activity" if act else "ca-carbonate-unit-activity"
    t, setup = ca_carbonate(act)
    cells, _ = speciate(t, setup)
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    assert abs(-np.log10(g[2] * m[2]) - gold(golds, name, "GENERIC: pH")) < TOL
    for j, sp in enumerate(["HCO3-", "Ca++", "H+"]):
        assert abs(cells["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < TOL
        if act:
            assert abs(g[j] - gold(golds, name, f"GENERIC: Gamma {sp}")) < TOL


def test_valocchi_ion_exchange(golds):
    name = "ion-exchange-valocchi"
    t, setup = valocchi()
    cells, _ = speciate(t, setup)
    for j, sp in enumerate(["Na+", "Ca++", "Mg++", "Cl-"]):
        assert abs(cells["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < TOL
        assert abs(cells["total_sorb"][0][j] - gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}")) < TOL
    # charge invariant of the exchanger: sum z_i S_i = CEC
    Z = t.Z
    assert abs(np.dot(Z, cells["total_sorb"][0]) - 750.0) < 1e-10


def test_sorbed_total_constraints_recover_the_valocchi_state(golds):
    """Sorbed-total constraint rows (reaction.F90:1490-1492, 1608-1614): with
    Ca++ and Mg++ constrained by the gold's sorbed totals, the solution is the
    gold's state again (aqueous totals, and the sorbed totals themselves);
    the sorbed rows' Jacobian is the exchanger's dS/dm, so a wrong row would
    stall or land elsewhere."""
    name = "ion-exchange-valocchi"
    sorbed = [gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}") for sp in ("Ca++", "Mg++")]
    t, setup = valocchi_sorbed(sorbed)
    cells, _ = speciate(t, setup)
    for j, sp in enumerate(["Na+", "Ca++", "Mg++", "Cl-"]):
        g_t = gold(golds, name, f"CONCENTRATION: Total {sp}")
        assert abs(cells["total"][0][j] - g_t) <= 1e-9 * abs(g_t)
        g_s = gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}")
        assert abs(cells["total_sorb"][0][j] - g_s) <= 1e-9 * max(abs(g_s), 1.0)


@pytest.mark.parametrize("activity", [False, True])
def test_mineral_and_gas_constraints(activity):
    """Mineral and gas constraint rows (reaction.F90:1694-1745) on the calcite
    decks' initial constraint.  With unit activity the solution has
    ln(Q/K) = 0 for calcite and ln(Q/K) = ln(1e-3 bar) for CO2(g) (from the
    database's log K, independently of the oracle); with activity on, the
    coefficients of the solve are not those of the refresh that follows, so
    the pin there is the inverse: speciating again from the solution's totals
    (TOTAL rows, H+ free) returns the same free ions."""
    from chemdefs import CON_FREE, CON_TOTAL, calcite_column_initial
    t, setup = calcite_column_initial(activity)
    cells, _ = speciate(t, setup)
    m = cells["pri_molal"][0]
    assert abs(m[0] - 1e-8 * 1000.0 / DEN) <= 1e-20
    if not activity:
        # hanford.dat at 25 C: Calcite = -H+ + Ca++ + HCO3-, log K 1.8487;
        # CO2(g) = -H2O + H+ + HCO3-, log K -7.8136
        la = np.log(m)
        assert abs(-1.8487 * 2.30258509299 - la[0] + la[2] + la[1]) <= 1e-10
        assert abs(7.8136 * 2.30258509299 + la[0] + la[1] - np.log(1e-3)) <= 1e-10
    again = dict(setup, ctype=[CON_FREE, CON_TOTAL, CON_TOTAL],
                 conc=[setup["conc"][0], cells["total"][0][1], cells["total"][0][2]])
    again.pop("rxn")
    cells2, _ = speciate(t, again)
    # the activity-coupled stop (|dm/m| < 1e-6 one iteration after activity
    # is on) leaves both solutions within a few 1e-9 of the fixed point
    assert np.max(np.abs(cells2["pri_molal"][0] - m) / m) <= (1e-8 if activity else 1e-10)


def test_calcite_kinetics(golds):
    name = "calcite-kinetics"
    t, setup = calcite_kinetics()
    cells, o = speciate(t, setup)
    o.rel_tol, o.extra_iterations, o.max_iterations = 1e-14, 1, 60
    orc.batch_steps(t, cells, o, [0.5] * 500)
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    assert abs(-np.log10(g[0] * m[0]) - gold(golds, name, "GENERIC: pH")) < TOL
    for j, sp in enumerate(["H+", "HCO3-", "Ca++"]):
        assert abs(cells["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < TOL
    assert abs(cells["mnrl_volfrac"][0, 0] - gold(golds, name, "VOLUME_FRACTION: Calcite VF")) < TOL
    assert abs(cells["mnrl_rate"][0, 0] - gold(golds, name, "RATE: Calcite Rate")) < TOL
