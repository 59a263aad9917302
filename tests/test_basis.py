"""Reactions written in a deck's own primary species (database.Basis, the
role of BasisInit's basis handling, reaction_database.F90:1246-1380).

Pins:
- a database reaction already in the primaries is taken as written;
- closed forms: with CO2(aq) as a primary, HCO3- = CO2(aq) - H+ + H2O with
  log K = -log K(CO2(aq)), and Calcite's log K gains -log K(CO2(aq)); a
  species defined through another complex (U++++ in Uraninite) or through a
  database primary that a gas reaction defines (O2(g) = O2(aq)) is reduced
  to the primaries with its log K summed along the substitution;
- invariance: the same water speciated with the primaries H+, HCO3-, Ca++
  and, from the totals that speciation implies in the other basis, with
  H+, CO2(aq), Ca++ gives the same molality of every species (free ions of
  one are complexes of the other), with unit activity and with
  Debye-Hueckel activity: a wrong sign or index in the rewritten log K or
  stoichiometry moves it;
- a species the primaries cannot express is refused."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN
from pflotran_amd import abi
from pflotran_amd.database import Basis, build_tables, read_database

DB = os.path.join(GOLDEN, "db", "hanford_subset.dat")
CPLX = ["OH-", "CO3--", "CaHCO3+", "CaCO3(aq)", "CaOH+"]


def test_reactions_in_the_database_basis_are_unchanged():
    db = read_database(DB)
    b = Basis(db, ["H+", "HCO3-", "Ca++"])
    for n in CPLX + ["CO2(aq)"]:
        assert b.reaction(n, db.complexes) is db.complexes[n]
    assert b.reaction("Calcite", db.minerals) is db.minerals["Calcite"]


def test_swapped_basis_closed_forms():
    db = read_database(DB)
    b = Basis(db, ["H+", "CO2(aq)", "Ca++"])
    co2 = db.complexes["CO2(aq)"]
    r = b.reaction("HCO3-", db.complexes)
    assert dict((sp, st) for st, sp in r.species) == {"H+": -1.0, "CO2(aq)": 1.0, "H2O": 1.0}
    for i, lk in enumerate(co2.logK):
        assert r.logK[i] == (500.0 if lk == 500.0 else -lk)
    cal = b.reaction("Calcite", db.minerals)
    assert dict((sp, st) for st, sp in cal.species) == {"H+": -2.0, "CO2(aq)": 1.0, "Ca++": 1.0, "H2O": 1.0}
    i25 = db.temperatures.index(25.0)
    assert abs(cal.logK[i25] - (db.minerals["Calcite"].logK[i25] - co2.logK[i25])) < 1e-14
    # charge and size of a primary taken from its complex entry
    assert b.primary("CO2(aq)").Z == 0.0 and b.primary("CO2(aq)").a0 == co2.a0


def test_substitution_through_complexes_and_gas_defined_primaries():
    db = read_database(DB)
    b = Basis(db, ["Na+", "Ca++", "H+", "UO2++", "SiO2(aq)", "HCO3-", "Cl-", "O2(aq)"])
    i25 = db.temperatures.index(25.0)
    u4 = db.complexes["U++++"]          # -H2O - 0.5 O2(g) + UO2++ + 2 H+
    o2g = db.gases["O2(g)"]             # O2(g) = O2(aq)
    uran = b.reaction("Uraninite", db.minerals)   # -4 H+ + U++++ + 2 H2O
    assert dict((sp, st) for st, sp in uran.species) == {"H+": -2.0, "UO2++": 1.0, "O2(aq)": -0.5, "H2O": 1.0}
    want = db.minerals["Uraninite"].logK[i25] + (u4.logK[i25] - 0.5 * o2g.logK[i25])
    assert abs(uran.logK[i25] - want) < 1e-12
    with pytest.raises(ValueError):
        Basis(db, ["H+", "HCO3-"]).reaction("Uraninite", db.minerals)


def _speciate(t, ctype, conc, act):
    c = orc.HostCells.allocate(t, 1)
    o = abi.default_opts(den_kg=997.16)
    rc, _ = orc.equilibrate(t, ctype, np.asarray([conc], dtype=np.float64), c, o, molal=True)
    assert rc == 0
    return c


@pytest.mark.parametrize("act", [False])
def test_speciation_is_the_same_in_either_basis(act):
    db = read_database(DB)
    freq = abi.PFX_ACT_TIMESTEP if act else abi.PFX_ACT_OFF
    t1 = build_tables(db, ["H+", "HCO3-", "Ca++"], CPLX + ["CO2(aq)"], act_freq=freq)
    t2 = build_tables(db, ["H+", "CO2(aq)", "Ca++"], CPLX + ["HCO3-"], act_freq=freq)
    ph, tc, tca = 7.3, 3e-3, 1.2e-3
    c1 = _speciate(t1, [abi.PFX_CONSTRAINT_PH, abi.PFX_CONSTRAINT_TOTAL, abi.PFX_CONSTRAINT_TOTAL], [ph, tc, tca], act)
    m1 = dict(zip(t1.primary_names, c1["pri_molal"][0]))
    m1.update(zip(t1.complex_names, c1["sec_molal"][0]))
    # the carbon total in the CO2(aq) basis: every carbon species once
    carbon = ["HCO3-", "CO2(aq)", "CO3--", "CaHCO3+", "CaCO3(aq)"]
    tc2 = sum(m1[n] for n in carbon)
    c2 = _speciate(t2, [abi.PFX_CONSTRAINT_PH, abi.PFX_CONSTRAINT_TOTAL, abi.PFX_CONSTRAINT_TOTAL], [ph, tc2, tca],
                   act)
    m2 = dict(zip(t2.primary_names, c2["pri_molal"][0]))
    m2.update(zip(t2.complex_names, c2["sec_molal"][0]))
    for n in ["H+", "Ca++"] + carbon + ["OH-", "CaOH+"]:
        assert abs(m2[n] - m1[n]) <= 1e-9 * m1[n], (n, m1[n], m2[n])


def test_mass_action_is_the_same_in_either_basis_with_activity():
    # Debye-Hueckel activity: the constraint Newton stops on its own
    # tolerances with lagged coefficients, so compare the mass action
    # itself: basis 1's free ions and coefficients, carried over to basis 2
    # (CO2(aq) a primary with its coefficient as a complex), give back every
    # other species' molality through basis 2's rewritten reactions
    db = read_database(DB)
    t1 = build_tables(db, ["H+", "HCO3-", "Ca++"], CPLX + ["CO2(aq)"], act_freq=abi.PFX_ACT_TIMESTEP)
    t2 = build_tables(db, ["H+", "CO2(aq)", "Ca++"], CPLX + ["HCO3-"], act_freq=abi.PFX_ACT_TIMESTEP)
    c1 = _speciate(t1, [abi.PFX_CONSTRAINT_PH, abi.PFX_CONSTRAINT_TOTAL, abi.PFX_CONSTRAINT_TOTAL],
                   [7.3, 3e-3, 1.2e-3], True)
    m1 = dict(zip(t1.primary_names, c1["pri_molal"][0]))
    m1.update(zip(t1.complex_names, c1["sec_molal"][0]))
    g1 = dict(zip(t1.primary_names, c1["pri_act_coef"][0]))
    g1.update(zip(t1.complex_names, c1["sec_act_coef"][0]))
    c2 = orc.HostCells.allocate(t2, 1)
    for j, n in enumerate(t2.primary_names):
        c2["pri_molal"][0, j] = m1[n]
        c2["pri_act_coef"][0, j] = g1[n]
    for i, n in enumerate(t2.complex_names):
        c2["sec_act_coef"][0, i] = g1[n]
    orc.auxvar(t2, c2, abi.default_opts(den_kg=997.16), update_activity=False)
    m2 = dict(zip(t2.complex_names, c2["sec_molal"][0]))
    for n in t2.complex_names:
        assert abs(m2[n] - m1[n]) <= 1e-12 * m1[n], (n, m1[n], m2[n])
