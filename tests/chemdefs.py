"""Chemistry of the reference's batch regression tests
(regression_tests/ascem/batch/*.in), restated as table builds."""
import os

from pflotran_amd import abi
from pflotran_amd.database import (IonExchangeRxn, MineralKinetics, build_tables, constraint_rxn, constraint_rxns,
                                   read_database)

DB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "db")

CON_TOTAL, CON_FREE, CON_LOG, CON_PH, CON_CHARGE, CON_TOTAL_SORB, CON_MINERAL, CON_GAS = 0, 1, 2, 3, 4, 5, 6, 7


def carbonate(activity):
    """carbonate-{unit,debye-huckel}-activity.in: H+/HCO3- totals 1e-3, porosity 0.25"""
    db = read_database(os.path.join(DB, "carbonate.dat"))
    t = build_tables(db, ["H+", "HCO3-"], ["OH-", "CO3--", "CO2(aq)"],
                     act_freq=abi.PFX_ACT_TIMESTEP if activity else abi.PFX_ACT_OFF)
    return t, dict(ctype=[CON_TOTAL, CON_TOTAL], conc=[1e-3, 1e-3], porosity=0.25)


def ca_carbonate(activity):
    """ca-carbonate-{unit,debye-huckel}-activity.in"""
    db = read_database(os.path.join(DB, "ca-carbonate.dat"))
    t = build_tables(db, ["HCO3-", "Ca++", "H+"], ["OH-", "CO2(aq)", "CO3--", "CaOH+", "CaHCO3+", "CaCO3(aq)"],
                     act_freq=abi.PFX_ACT_TIMESTEP if activity else abi.PFX_ACT_OFF)
    return t, dict(ctype=[CON_TOTAL] * 3, conc=[3e-3, 1e-3, 1e-3], porosity=0.25)


def calcite_kinetics(vf_test=False):
    """calcite-kinetics.in (500 x 0.5 s) and calcite-kinetics-volume-fractions.in
    (30 d steps to 60 months); calcite k = 1e-13 mol/cm^2/s = 1e-9 mol/m^2/s,
    vf 0.2, area 1 cm^2/cm^3 = 100 m^2/m^3, porosity 0.5"""
    db = read_database(os.path.join(DB, "calcite.dat"))
    t = build_tables(db, ["H+", "HCO3-", "Ca++"], ["OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"],
                     {"Calcite": MineralKinetics(1e-9)}, act_freq=abi.PFX_ACT_TIMESTEP)
    conc = [1e-2, 1e-2, 1e-19] if vf_test else [-1e-5, 1e-5, 1e-5]
    return t, dict(ctype=[CON_TOTAL] * 3, conc=conc, porosity=0.5, vf=0.2, area=100.0)


def valocchi():
    """ion-exchange-valocchi.in: Na+/Ca++/Mg++ totals, Cl- charge balance,
    CEC 750 eq/m^3 with Na+ as reference cation; halite (k = 1e-40 mol/cm^2/s)"""
    db = read_database(os.path.join(DB, "hanford_subset.dat"))
    t = build_tables(db, ["Na+", "Ca++", "Mg++", "Cl-"], [], {"Halite": MineralKinetics(1e-36)},
                     [IonExchangeRxn([("Na+", 1.0), ("Ca++", 3.38638672536), ("Mg++", 6.00240096038)], 750.0)])
    return t, dict(ctype=[CON_TOTAL, CON_TOTAL, CON_TOTAL, CON_CHARGE], conc=[8.65e-2, 1.82e-2, 1.11e-2, 2e-3],
                   porosity=0.25, vf=1e-5, area=100.0)


def valocchi_sorbed(golds_sorbed):
    """ion-exchange-valocchi with Ca++ and Mg++ constrained by their sorbed
    totals (mol/m^3 bulk; the reference's TOTAL_SORB constraint) instead of
    their aqueous totals.  The three sorbed totals alone are not independent
    (sum z S = CEC), so Na+ keeps its aqueous total."""
    t, setup = valocchi()
    setup = dict(setup, ctype=[CON_TOTAL, CON_TOTAL_SORB, CON_TOTAL_SORB, CON_CHARGE],
                 conc=[setup["conc"][0], golds_sorbed[0], golds_sorbed[1], setup["conc"][3]])
    return t, setup


def calcite_column_initial(activity=True):
    """The initial constraint of the 1-D calcite transport decks
    (shortcourse/1D_Calcite/calcite_tran_only.in, ascem/1d/1d-calcite):
    H+ 1e-8 free, HCO3- in equilibrium with CO2(g) at 1e-3 bar, Ca++ in
    equilibrium with calcite (5e-4 the starting value), porosity 0.25."""
    db = read_database(os.path.join(DB, "hanford_subset.dat"))
    prim = ["H+", "HCO3-", "Ca++"]
    t = build_tables(db, prim, ["OH-", "CO3--", "CO2(aq)", "CaCO3(aq)", "CaHCO3+", "CaOH+"],
                     act_freq=abi.PFX_ACT_TIMESTEP if activity else abi.PFX_ACT_OFF)
    rxn = constraint_rxns(t, {1: constraint_rxn(db, "CO2(g)", prim, "gas"),
                              2: constraint_rxn(db, "Calcite", prim, "mineral")})
    return t, dict(ctype=[CON_FREE, CON_GAS, CON_MINERAL], conc=[1e-8, 1e-3, 5e-4], porosity=0.25, rxn=rxn)


def carbonate_column():
    """The 1-D calcite column chemistry: 5 primaries, 8 complexes and kinetic
    calcite (k = 1e-9 mol/m^2/s) from the hanford database, 25 C."""
    db = read_database(os.path.join(DB, "hanford_subset.dat"))
    t = build_tables(db, ["H+", "HCO3-", "Ca++", "Mg++", "Cl-"],
                     ["OH-", "CO2(aq)", "CO3--", "CaHCO3+", "CaCO3(aq)", "MgHCO3+", "MgCO3(aq)", "CaCl+"],
                     {"Calcite": MineralKinetics(1e-9)}, act_freq=abi.PFX_ACT_TIMESTEP)
    return t
