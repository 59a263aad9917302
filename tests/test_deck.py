"""The deck reader (pflotran_amd/deck.py) on the reference's own input decks
(regression_tests/ascem/batch/*.in, ascem/1d/1d-calcite/1d-calcite.in,
copied under tests/golden/decks): the tables it builds equal the ones the
chemistry restatements build by hand, and the oracle fed from the decks
lands on the regression golds (1e-12 absolute, batch.cfg)."""
import os

import numpy as np
import pytest

import oracle as orc
from chemdefs import DB, ca_carbonate, calcite_kinetics, carbonate, valocchi
from conftest import GOLDEN, gold
from pflotran_amd import abi
from pflotran_amd.database import water_density_ifc67
from pflotran_amd.deck import DeckError, parse_deck, read_deck, units_factor

DECKS = os.path.join(GOLDEN, "decks")
DEN = water_density_ifc67()
TOL = 1e-12
# the decks name the reference's database files; the fixtures hold them
# (hanford.dat as an excerpt of the species these decks use)
DB_FILES = {"hanford.dat": "hanford_subset.dat"}


def load(name):
    d = read_deck(os.path.join(DECKS, f"{name}.in"))
    base = os.path.basename(d.chemistry.database)
    db = d.database(os.path.join(DB, DB_FILES.get(base, base)))
    return d, db


def speciate(d, db, constraint="initial"):
    t = d.tables(db)
    s = d.constraint_setup(constraint, t, db)
    cells = orc.HostCells.allocate(t, 1)
    if s["vf"] is not None:
        cells["mnrl_volfrac"][0] = s["vf"]
        cells["mnrl_area"][0] = s["area"]
    o = abi.default_opts(porosity=d.porosity["soil1"], den_kg=DEN)
    rc, its = orc.equilibrate(t, s["ctype"], np.array(s["conc"]), cells, o, molal=s["molal"], rxn=s["rxn"])
    assert rc == 0 and its[0] > 0
    return t, cells, o


TABLE_FIELDS = ("Z", "a0", "cplx_specid", "cplx_stoich", "cplx_h2ostoich", "cplx_Z", "cplx_a0", "cplx_logK",
                "mnrl_specid", "mnrl_stoich", "mnrl_h2ostoich", "mnrl_logK",
                "mnrl_molar_vol", "mnrl_activation_energy", "ionx_cationid", "ionx_k", "ionx_CEC")


@pytest.mark.parametrize("name,by_hand", [
    ("carbonate-unit-activity", lambda: carbonate(False)),
    ("carbonate-debye-huckel-activity", lambda: carbonate(True)),
    ("ca-carbonate-unit-activity", lambda: ca_carbonate(False)),
    ("ca-carbonate-debye-huckel-activity", lambda: ca_carbonate(True)),
    ("calcite-kinetics", lambda: calcite_kinetics()),
    ("calcite-kinetics-volume-fractions", lambda: calcite_kinetics(vf_test=True)),
    ("ion-exchange-valocchi", valocchi),
])
def test_deck_tables_match_the_restated_chemistry(name, by_hand):
    d, db = load(name)
    t = d.tables(db)
    ref, setup = by_hand()
    assert t.primary_names == ref.primary_names and t.complex_names == ref.complex_names
    assert t.mineral_names == ref.mineral_names
    for f in TABLE_FIELDS:
        a, b = getattr(t, f), getattr(ref, f)
        assert a.shape == b.shape and np.array_equal(a, b), f
    # unit conversions (1e-13 mol/cm^2/s, 1 cm^2/cm^3) agree to rounding
    assert np.allclose(t.mnrl_rate_constant, ref.mnrl_rate_constant, rtol=1e-15, atol=0)
    assert (t.act_freq, t.act_alg, t.use_log_formulation, t.debye) == \
        (ref.act_freq, ref.act_alg, ref.use_log_formulation, ref.debye)
    s = d.constraint_setup("initial", t, db)
    assert s["ctype"] == setup["ctype"] and np.array_equal(s["conc"], setup["conc"])
    assert d.porosity["soil1"] == setup["porosity"]
    if "vf" in setup:
        assert s["vf"][0] == setup["vf"] and s["area"][0] == pytest.approx(setup["area"], rel=1e-15)


@pytest.mark.parametrize("name,ph_col", [
    ("carbonate-unit-activity", 0), ("carbonate-debye-huckel-activity", 0),
    ("ca-carbonate-unit-activity", 2), ("ca-carbonate-debye-huckel-activity", 2),
    ("ion-exchange-valocchi", None)])
def test_deck_speciation_golds(golds, name, ph_col):
    d, db = load(name)
    t, cells, _ = speciate(d, db)
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    if ph_col is not None:
        assert abs(-np.log10(g[ph_col] * m[ph_col]) - gold(golds, name, "GENERIC: pH")) < TOL
    for j, sp in enumerate(t.primary_names):
        assert abs(cells["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < TOL
        if t.R:
            assert abs(cells["total_sorb"][0][j] - gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}")) < TOL


def test_deck_calcite_kinetics_gold(golds):
    name = "calcite-kinetics"
    d, db = load(name)
    steps = d.time_steps()
    assert len(steps) == 500 and set(steps) == {0.5}
    t, cells, o = speciate(d, db)
    o.rel_tol, o.extra_iterations, o.max_iterations = 1e-14, 1, 60
    orc.batch_steps(t, cells, o, steps)
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    assert abs(-np.log10(g[0] * m[0]) - gold(golds, name, "GENERIC: pH")) < TOL
    for j, sp in enumerate(t.primary_names):
        assert abs(cells["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < TOL
    assert abs(cells["mnrl_volfrac"][0, 0] - gold(golds, name, "VOLUME_FRACTION: Calcite VF")) < TOL
    assert abs(cells["mnrl_rate"][0, 0] - gold(golds, name, "RATE: Calcite Rate")) < TOL


def test_deck_time_schedule_lands_on_the_final_time():
    d, _ = load("calcite-kinetics-volume-fractions")
    steps = d.time_steps()
    # 60 months of 365/12 days in 30-day steps: 60 full steps and 25 days
    assert len(steps) == 61 and steps[:60] == [30 * 86400.0] * 60
    assert abs(steps[-1] - 25 * 86400.0) < 1e-6 and abs(sum(steps) - 60 * 365 / 12 * 86400.0) < 1e-3


def test_deck_1d_calcite_constraints():
    """The 1-D calcite deck's two constraints: 'initial' (H+ free, HCO3-
    from CO2(g) at 1e-3 bar, Ca++ from calcite) and 'west' (pH 5, HCO3-
    total, Ca++ by charge balance).  Pins: the free and pH rows hold exactly,
    the total row holds, the charge-balance row zeroes the charge of the
    totals, and the gas and mineral rows hold in their database log K
    (hanford.dat at 25 C: Calcite log K 1.8487, CO2(g) log K -7.8136)."""
    d, db = load("1d-calcite")
    assert d.chemistry.gases == ["CO2(g)"] and d.kinetic_minerals() == ["Calcite"]
    assert d.chemistry.kinetics["Calcite"].rate_constant == pytest.approx(1e-9, rel=1e-15)
    t, cells, _ = speciate(d, db, "initial")
    m, g = cells["pri_molal"][0], cells["pri_act_coef"][0]
    assert abs(m[0] - 1e-8 * 1000.0 / DEN) <= 1e-20
    s = d.constraint_setup("initial", t, db)
    assert s["vf"][0] == 1e-5 and s["area"][0] == pytest.approx(100.0, rel=1e-15)
    t2, cells2, _ = speciate(d, db, "west")
    m2, g2 = cells2["pri_molal"][0], cells2["pri_act_coef"][0]
    # activity is on: the rows hold for the coefficients of the last solve
    # and the refresh that follows moves them by the activity-coupled stop
    # (|dm/m| < 1e-6 one iteration after activity is on), a few 1e-12 here
    assert abs(-np.log10(g2[0] * m2[0]) - 5.0) <= 1e-10
    assert abs(cells2["total"][0][1] - 1e-3) <= 1e-13
    assert abs(np.dot(t2.Z, cells2["total"][0])) <= 1e-10 * np.max(np.abs(cells2["total"][0]))


def test_units():
    assert units_factor("mol/cm^2-sec", "mol/m^2-sec") == pytest.approx(1e4, rel=1e-15)
    assert units_factor("cm^2/cm^3", "m^2/m^3") == pytest.approx(100.0, rel=1e-15)
    assert units_factor("kJ/mol", "J/mol") == 1e3
    assert units_factor("mo", "s") == 365.0 / 12.0 * 86400.0
    assert units_factor("y", "d") == 365.0
    with pytest.raises(DeckError):
        units_factor("furlong", "m")


MINI = """
# comment line
SIMULATION
  PROCESS_MODELS
    SUBSURFACE_TRANSPORT transport
      GLOBAL_IMPLICIT
      NUMERICAL_JACOBIAN
    /
  /
END
CHEMISTRY
  PRIMARY_SPECIES
    Na+   ! trailing comment
    Ca++
  /
SKIP
  SECONDARY_SPECIES
    junk
  /
NOSKIP
  SORPTION
    ION_EXCHANGE_RXN
      CEC 10. eq/m^3
      CATIONS
        Ca++ 2.5
        Na+ 1.d0 REFERENCE
      /
    /
  /
  DATABASE x.dat
  ACTIVITY_COEFFICIENTS NEWTON NEWTON_ITERATION
  MAX_DLNC 2.5
  MOLAL
  OUTPUT
    TOTAL
  /
END
CONSTRAINT c1
  CONCENTRATIONS
    Ca++ 1.d-3
    Na+  7. P
  /
END
MATERIAL_PROPERTY soil1
  PERMEABILITY
    PERM_ISO 1.d-12
  /
  POROSITY 0.3
END
TIME
  FINAL_TIME 1 h
  INITIAL_TIMESTEP_SIZE 1 min
END
"""


def test_parser_rules():
    d = parse_deck(MINI)
    c = d.chemistry
    assert c.primaries == ["Na+", "Ca++"] and c.secondaries == []
    assert d.numerical_jacobian and c.molal and c.max_dlnC == 2.5
    assert (c.act_alg, c.act_freq) == (abi.PFX_ACT_ALG_NEWTON, abi.PFX_ACT_NEWTON_ITER)
    # the reference cation moves to the front (reaction.F90:700-721)
    assert c.ion_exchange[0].cations == [("Na+", 1.0), ("Ca++", 2.5)] and c.ion_exchange[0].CEC == 10.0
    con = d.constraints["c1"]
    assert [(s.name, s.value, s.ctype) for s in con.concentrations] == \
        [("Ca++", 1e-3, abi.PFX_CONSTRAINT_TOTAL), ("Na+", 7.0, abi.PFX_CONSTRAINT_PH)]
    assert d.porosity == {"soil1": 0.3}
    assert d.time_steps() == [60.0] * 60


@pytest.mark.parametrize("edit,message", [
    (("Na+ 1.d0 REFERENCE", "Na+ 1.d0"), "Reference cation missing"),
    (("Na+ 1.d0 REFERENCE", "Na+ 2.d0 REFERENCE"), "must have k = 1"),
    (("Na+  7. P", "Na+  7. S"), "no longer supported"),
    (("MAX_DLNC 2.5", "UPDATE_POROSITY"), "outside the per-cell path"),
    (("MAX_DLNC 2.5", "FOO"), "unknown"),
    (("    Ca++ 1.d-3\n", ""), None),
])
def test_parser_errors(edit, message):
    d = None
    with pytest.raises(DeckError, match=message):
        d = parse_deck(MINI.replace(*edit))
        # a constraint that misses a primary fails where its arrays are built
        d.constraint_setup("c1", None, None)


def test_external_file_and_waypoint_times(tmp_path):
    """EXTERNAL_FILE lines are replaced by the named file's lines
    (input_aux.F90:682-686); a TIME card with a waypoint maximum parses, and
    only the schedule, which does not model waypoints, refuses it."""
    (tmp_path / "chem.txt").write_text("CHEMISTRY\n  PRIMARY_SPECIES\n    H+\n    HCO3-\n  /\n  DATABASE x.dat\nEND\n")
    (tmp_path / "main.in").write_text(
        "EXTERNAL_FILE chem.txt\nCONSTRAINT c\n  CONCENTRATIONS\n    H+ 7. P\n    HCO3- 1.d-3 T\n  /\nEND\n"
        "TIME\n  FINAL_TIME 1 d\n  INITIAL_TIMESTEP_SIZE 1 h\n  MAXIMUM_TIMESTEP_SIZE 1 h\n"
        "  MAXIMUM_TIMESTEP_SIZE 6 h at 12 h\nEND\n")
    d = read_deck(str(tmp_path / "main.in"))
    assert d.chemistry.primaries == ["H+", "HCO3-"] and list(d.constraints) == ["c"]
    with pytest.raises(DeckError, match="waypoints"):
        d.time_steps()
    with pytest.raises(DeckError, match="EXTERNAL_FILE"):
        parse_deck((tmp_path / "main.in").read_text())


def test_tracer_deck_without_database():
    """A deck with primaries only and no DATABASE card: the reference skips
    full geochemistry (reaction.F90:899-906), turns the activity update off
    (:920-921) and takes a total as molarity, so the free ion is
    T 1000 / rho_w (:3376-3385).  Any reaction without a database is an
    error."""
    text = ("CHEMISTRY\n  PRIMARY_SPECIES\n    Tracer\n    Tracer2\n  /\n  OUTPUT\n    ALL\n  /\nEND\n"
            "CONSTRAINT initial\n  CONCENTRATIONS\n    Tracer 1.d-3 T\n    Tracer2 2.5d-7 T\n  /\nEND\n")
    d = parse_deck(text)
    assert d.tracer_only()
    db = d.database()
    t = d.tables(db)
    assert t.act_freq == abi.PFX_ACT_OFF and np.all(np.asarray(t.Z) == 0.0)
    s = d.constraint_setup("initial", t, db)
    cells = orc.HostCells.allocate(t, 1)
    o = abi.default_opts(den_kg=DEN)
    rc, _ = orc.equilibrate(t, s["ctype"], np.array(s["conc"]), cells, o, molal=s["molal"], rxn=s["rxn"])
    assert rc == 0
    assert np.allclose(cells["pri_molal"][0], np.array([1e-3, 2.5e-7]) * 1000.0 / DEN, rtol=1e-13, atol=0)
    assert np.allclose(cells["total"][0], [1e-3, 2.5e-7], rtol=1e-13, atol=0)
    assert np.all(cells["pri_act_coef"][0] == 1.0)
    with_mineral = text.replace("  OUTPUT", "  MINERALS\n    Calcite\n  /\n  OUTPUT")
    with pytest.raises(DeckError, match="no DATABASE"):
        parse_deck(with_mineral).database()


def test_time_step_growth():
    """TS_ACCELERATION and the default tfac (timestepper_BE.F90:124-133)
    with PMRTUpdateTimestep's rule (pm_rt.F90:631-647), pinned by closed
    forms: a step converged in 1-5 iterations doubles dt, up to the maximum;
    the schedule then lands on the final time.  The exact sequence of a
    reference run also depends on its Newton counts, which the caller
    supplies; no reference output records them, so the growth itself is
    parity unpinned beyond these rules."""
    d, _ = load("ca-carbonate-debye-huckel-activity")   # TS_ACCELERATION 8, 1e-6 y .. 0.1 y, 1 y
    assert d.time.iaccel == 8
    y = units_factor("y", "s")
    steps = d.time_steps()
    k = int(np.ceil(np.log2(0.1 / 1e-6)))            # doublings before the cap
    assert steps[:k] == pytest.approx([1e-6 * y * 2 ** i for i in range(k)], rel=1e-14)
    assert all(s == pytest.approx(0.1 * y, rel=1e-14) for s in steps[k:-1])
    assert sum(steps) == pytest.approx(1.0 * y, rel=1e-14) and steps[-1] > 0.1 * 0.1 * y
    # iteration counts beyond 5: tfac 1.8, 1.6, 1.4 up to TS_ACCELERATION 8, halving beyond
    dt = 1000.0
    assert [d.next_dt(dt, n) for n in (1, 5, 6, 7, 8, 9, 20)] == pytest.approx(
        [2000.0, 2000.0, 1800.0, 1600.0, 1400.0, 500.0, 500.0], rel=1e-15)
    assert d.next_dt(0.09 * y, 1) == pytest.approx(0.1 * y, rel=1e-15)   # capped at the maximum
    # the default TS_ACCELERATION (5): a 6-iteration step halves dt
    d2 = parse_deck("CHEMISTRY\n  PRIMARY_SPECIES\n    A\n  /\nEND\nTIME\n  FINAL_TIME 10 s\n"
                    "  INITIAL_TIMESTEP_SIZE 1 s\n  MAXIMUM_TIMESTEP_SIZE 4 s\nEND\n"
                    "TIMESTEPPER TRANSPORT\n  DT_FACTOR 1.5 1.5 1.5\nEND\n")
    assert d2.time.iaccel == 5 and d2.time.tfac == (1.5, 1.5, 1.5)
    assert d2.next_dt(2.0, 3) == 3.0 and d2.next_dt(2.0, 4) == 1.0 and d2.next_dt(2.0, 6) == 1.0
    # steps of 1, 1.5, 2.25, 3.375: the next would pass 10 s within 0.1 of a step
    assert d2.time_steps() == pytest.approx([1.0, 1.5, 2.25, 3.375, 10.0 - 8.125], rel=1e-15)
