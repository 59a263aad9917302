"""Shared by the run-level gold tests: the decks whose golds come from a
global-implicit run (the 1-cell batch kinetics decks and the 1-D columns),
their databases, and the gold's variables computed from a final state."""
import json
import os

import numpy as np

from conftest import GOLDEN

# deck -> (database fixture, gold file)
RUN_DECKS = {
    "calcite-kinetics": ("calcite.dat", "batch_golds.json"),
    "calcite-kinetics-volume-fractions": ("calcite.dat", "batch_golds.json"),
    "1d-calcite": ("hanford_subset.dat", "column_golds.json"),
    "calcite_tran_only": ("hanford_subset.dat", "column_golds.json"),
    "valocchi_ionx": ("hanford_subset.dat", "column_golds.json"),
    "solute_KD": ("hanford_subset.dat", "column_golds.json"),
    "solute_KD_w_mineral": ("hanford_subset.dat", "column_golds.json"),
    "solute_KD_wo_mineral": ("hanford_subset.dat", "column_golds.json"),
    "mineral_prefactor": ("hanford_subset.dat", "column_golds.json"),
}

# the suites' comparison criteria (regression_tests/*/*.cfg): 1e-12 absolute
# by default; default/column/column.cfg relaxes mineral_prefactor's
# concentrations to 1e-6 relative and its volume fractions to 5e-12 relative
CRITERIA = {"mineral_prefactor": {"CONCENTRATION": ("relative", 1e-6), "VOLUME_FRACTION": ("relative", 5e-12)}}


def within(name, key, x, v):
    """|x - gold| within the deck's criterion for the variable's kind.  One
    allowance: a volume fraction below 1e-12 grown by a mineral held within
    about 1e-9 of equilibrium (mineral_prefactor's Coffinite at cell 6,
    |1 - Q/K| ~ 6e-10 over the run) carries the cancellation of 1 - Q/K, a
    few ulps of ln(Q/K) being 1e-5 of the rate; it is held to 1e-18
    absolute instead of 5e-12 relative."""
    kind, tol = CRITERIA.get(name, {}).get(key.split(":")[0], ("absolute", 1e-12))
    if name in CRITERIA and key.startswith("VOLUME_FRACTION") and abs(v) < 1e-12:
        return abs(x - v) <= 1e-18
    return abs(x - v) <= (tol * abs(v) if kind == "relative" else tol)


def load(name):
    from pflotran_amd.database import read_database
    from pflotran_amd.deck import read_deck
    dbf, gf = RUN_DECKS[name]
    deck = read_deck(os.path.join(GOLDEN, "decks", f"{name}.in"))
    db = read_database(os.path.join(GOLDEN, "db", dbf))
    with open(os.path.join(GOLDEN, gf)) as f:
        g = json.load(f)[name]
    return deck, db, g


def outputs(t, state, den, molal=False):
    """The gold's variables per cell from the final state: pH from the
    activity of H+, totals in mol/L, free ions in mol/L (molality times the
    water density / 1000, as the reference prints them), or both in mol/kg
    water for a MOLAL deck (reaction.F90:878-891, patch.F90:5384-5387),
    sorbed totals in mol/m^3, volume fractions and rates."""
    names = t.primary_names
    out = {}
    if "H+" in names:
        h = names.index("H+")
        out["GENERIC: pH"] = -np.log10(state["pri_act_coef"][:, h] * state["pri_molal"][:, h])
    for j, sp in enumerate(names):
        if molal:
            out[f"CONCENTRATION: Total {sp}"] = state["total"][:, j] / den * 1000.0
            out[f"CONCENTRATION: Free {sp}"] = state["pri_molal"][:, j]
        else:
            out[f"CONCENTRATION: Total {sp}"] = state["total"][:, j]
            out[f"CONCENTRATION: Free {sp}"] = state["pri_molal"][:, j] * den * 1e-3
        if state.get("total_sorb") is not None:
            out[f"CONCENTRATION: Total Sorbed {sp}"] = state["total_sorb"][:, j]
    for k, mn in enumerate(t.mineral_names):
        out[f"VOLUME_FRACTION: {mn} VF"] = state["mnrl_volfrac"][:, k]
        out[f"RATE: {mn} Rate"] = state["mnrl_rate"][:, k]
    return out


def solution_norm(state):
    """The transport solution's 2-norm as the gold's SOLUTION block prints
    it: the solution vector holds every cell's free-ion molalities, also
    under the log formulation (which changes the update, not the vector)."""
    return float(np.linalg.norm(np.asarray(state["pri_molal"])))


def differences(g, out):
    """(|value - gold|, variable, cell, value, gold) for every gold value at
    a cell, largest first; at least one variable must be compared."""
    rows = []
    for key, cells in g.items():
        if key not in out:
            continue
        for cell, v in cells.items():
            if cell in ("Max", "Min", "Mean"):
                continue
            x = float(out[key][int(cell) - 1])
            rows.append((abs(x - v), key, cell, x, v))
    rows.sort(reverse=True)
    return rows
