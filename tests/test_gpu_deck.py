"""Decks through the device path: the tables and constraints the deck reader
builds from the reference's input decks, speciated by pfx_equilibrate_batch
on the GPU, against the oracle fed from the same deck and against the
regression golds (regression_tests/ascem/batch/*.regression.gold)."""
import numpy as np
import pytest

import oracle as orc
from conftest import gold
from pflotran_amd import abi, runtime, synth
from test_deck import DEN, load

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("name,constraint", [
    ("carbonate-unit-activity", "initial"), ("carbonate-debye-huckel-activity", "initial"),
    ("ca-carbonate-unit-activity", "initial"), ("ca-carbonate-debye-huckel-activity", "initial"),
    ("calcite-kinetics", "initial"), ("calcite-kinetics-volume-fractions", "initial"),
    ("ion-exchange-valocchi", "initial"), ("ion-exchange-valocchi", "inlet"),
    ("1d-calcite", "initial"), ("1d-calcite", "west")])
def test_gpu_deck_constraint(golds, name, constraint):
    d, db = load(name)
    t = d.tables(db)
    s = d.constraint_setup(constraint, t, db)
    n = 64
    # the deck's constraint in every cell, its concentrations scaled per
    # cell by up to +-20 % (pH and log rows shifted by up to +-0.1)
    rng = np.random.default_rng(5)
    conc = np.tile(np.array(s["conc"], dtype=np.float64), (n, 1))
    for j, ct in enumerate(s["ctype"]):
        if ct in (abi.PFX_CONSTRAINT_PH, abi.PFX_CONSTRAINT_LOG):
            conc[1:, j] += rng.uniform(-0.1, 0.1, n - 1)
        else:
            conc[1:, j] *= rng.uniform(0.8, 1.2, n - 1)
    o = synth.react_opts(porosity=d.porosity["soil1"])
    arr = synth.host_state_arrays(t, n, with_temperature=False)
    ref = orc.HostCells.allocate(t, n)
    if s["vf"] is not None:
        arr["mnrl_volfrac"][:] = s["vf"]
        arr["mnrl_area"][:] = s["area"]
        ref["mnrl_volfrac"][:] = s["vf"]
        ref["mnrl_area"][:] = s["area"]
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(arr, n)
    print(f"deck {name} constraint {constraint}: {n} cells", flush=True)
    rc, st = runtime.equilibrate(tab, dc, s["ctype"], conc, o, molal=bool(s["molal"]), rxn=s["rxn"])
    assert rc == 0, st.as_dict()
    # the initial-condition refresh the oracle's batch call includes: unit
    # coefficients and 1e-9 exchange state, then (activity on) two passes
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    if dc.t["ionx_ref_sorbed"] is not None:
        dc.t["ionx_ref_sorbed"].fill_(1e-9)
    runtime.auxvar(tab, dc, 2 if t.act_freq != abi.PFX_ACT_OFF else 0, o)
    out = dc.to_host()
    rc, _ = orc.equilibrate(t, s["ctype"], conc, ref, synth.react_opts(porosity=d.porosity["soil1"]),
                            molal=s["molal"], rxn=s["rxn"])
    assert rc == 0
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10
    assert rel_err(out["pri_act_coef"], ref["pri_act_coef"]) <= 1e-10
    assert np.max(np.abs(out["total"] - ref["total"])) <= 1e-10 * np.max(np.abs(ref["total"]))
    if constraint == "initial" and name in golds:
        # cell 0 holds the deck's own constraint: the gold's totals
        for j, sp in enumerate(t.primary_names):
            key = f"CONCENTRATION: Total {sp}"
            if name.startswith("calcite-kinetics"):
                continue  # those golds are after the run, not at the constraint
            assert abs(out["total"][0][j] - gold(golds, name, key)) < 1e-12, key
