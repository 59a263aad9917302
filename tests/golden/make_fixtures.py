"""Regenerate the committed fixtures under tests/golden/ from the reference
tree (run where the reference tree is present; the device runs never read
the reference).  Fixtures are data: database excerpts, the input decks and
the regression-gold values of the reference's own batch chemistry tests.

    python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
BATCH = os.path.join(REF, "regression_tests", "ascem", "batch")

# database files held by the reference's batch tests, copied verbatim
for name in ("carbonate.dat", "ca-carbonate.dat", "calcite.dat"):
    shutil.copy(os.path.join(BATCH, name), os.path.join(HERE, "db", name))

# input decks of those tests (data: the inputs whose outputs the golds
# hold)
os.makedirs(os.path.join(HERE, "decks"), exist_ok=True)
for name in ("carbonate-unit-activity", "carbonate-debye-huckel-activity",
             "ca-carbonate-unit-activity", "ca-carbonate-debye-huckel-activity",
             "calcite-kinetics", "calcite-kinetics-volume-fractions", "ion-exchange-valocchi"):
    shutil.copy(os.path.join(BATCH, f"{name}.in"), os.path.join(HERE, "decks", f"{name}.in"))
# global-implicit column decks whose golds pin the device time loop
COLUMNS = {"1d-calcite": os.path.join("ascem", "1d", "1d-calcite"),
           "calcite_tran_only": os.path.join("shortcourse", "1D_Calcite"),
           "valocchi_ionx": os.path.join("default", "column"),
           # isotherm sorption: the linear K_D column and its 1-cell decks
           "solute_KD": os.path.join("default", "column"),
           "solute_KD_w_mineral": os.path.join("default", "batch"),
           "solute_KD_wo_mineral": os.path.join("default", "batch"),
           # mineral rate prefactors, a uranium basis with U++++ and O2(g)
           # reached through other reactions
           "mineral_prefactor": os.path.join("default", "column")}
for name, sub in COLUMNS.items():
    shutil.copy(os.path.join(REF, "regression_tests", sub, f"{name}.in"), os.path.join(HERE, "decks", f"{name}.in"))

# excerpt of database/hanford.dat: the species of the 1-D calcite columns
# (5 primaries, 8 complexes, calcite; CaOH+ and the CO2(g) constraint of the
# calcite transport decks), of the Valocchi ion-exchange test, of the K_D
# decks (A(aq), A(s)) and of the mineral-prefactor column (its CHEMISTRY
# species lists, read from the deck)
KEEP = {"H2O", "H+", "HCO3-", "Ca++", "Mg++", "Cl-", "Na+", "CO2(g)", "CaOH+",
        "OH-", "CO2(aq)", "CO3--", "CaHCO3+", "CaCO3(aq)", "MgHCO3+", "MgCO3(aq)", "CaCl+",
        "Calcite", "Halite", "A(aq)", "A(s)"}


def chemistry_names(deck):
    """Species named in a deck's PRIMARY_SPECIES, SECONDARY_SPECIES,
    GAS_SPECIES and MINERALS lists."""
    names, block = set(), None
    for ln in open(deck):
        w = ln.split("#")[0].split()
        if not w:
            continue
        if w[0] in ("PRIMARY_SPECIES", "SECONDARY_SPECIES", "GAS_SPECIES", "MINERALS") and block is None:
            block = w[0]
        elif block and w[0] in ("/", "END"):
            block = None
        elif block:
            names.add(w[0])
    return names


KEEP |= chemistry_names(os.path.join(REF, "regression_tests", "default", "column", "mineral_prefactor.in"))
out = []
with open(os.path.join(REF, "database", "hanford.dat")) as f:
    for i, ln in enumerate(f):
        m = re.match(r"^'([^']*)'", ln)
        if i == 0 or (m and (m.group(1) in KEEP or m.group(1) == "null")):
            out.append(ln if ln.endswith("\n") else ln + "\n")
with open(os.path.join(HERE, "db", "hanford_subset.dat"), "w") as f:
    f.writelines(out)


def parse_gold(path):
    vals, sec = {}, None
    for ln in open(path):
        m = re.match(r"^-- (\w+): (.*) --$", ln.strip())
        if m:
            sec = f"{m.group(1)}: {m.group(2)}"
            vals[sec] = {}
            continue
        m = re.match(r"^\s*([\w\s\-\(\)]+):\s+(\S+)$", ln.rstrip())
        if m and sec:
            key = m.group(1).strip()
            try:
                vals[sec][key] = float(m.group(2))
            except ValueError:
                pass
    return vals


golds = {}
for test in ("carbonate-unit-activity", "carbonate-debye-huckel-activity",
             "ca-carbonate-unit-activity", "ca-carbonate-debye-huckel-activity",
             "calcite-kinetics", "calcite-kinetics-volume-fractions",
             "ion-exchange-valocchi"):
    golds[test] = parse_gold(os.path.join(BATCH, f"{test}.regression.gold"))
golds["_source"] = "regression_tests/ascem/batch/*.regression.gold (tolerance 1e-12 absolute, batch.cfg)"
with open(os.path.join(HERE, "batch_golds.json"), "w") as f:
    json.dump(golds, f, indent=1, sort_keys=True)

columns = {name: parse_gold(os.path.join(REF, "regression_tests", sub, f"{name}.regression.gold"))
           for name, sub in COLUMNS.items()}
columns["_source"] = ("regression_tests/{ascem/1d/1d-calcite,shortcourse/1D_Calcite,default/column,"
                      "default/batch}/*.regression.gold (tolerance 1e-12 absolute in their .cfg files)")
with open(os.path.join(HERE, "column_golds.json"), "w") as f:
    json.dump(columns, f, indent=1, sort_keys=True)
print("fixtures written to", HERE)
