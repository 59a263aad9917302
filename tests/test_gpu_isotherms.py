"""Isotherm (K_D) sorption on the device against the oracle (whose
RTotalSorbKD restatement the CPU isotherm tests pin): a 15-primary network
with ion exchange, three kinetic minerals and four isotherms (linear,
Langmuir, Freundlich, linear scaled by a mineral's volume fraction) through
the residual and Jacobian, the reaction step and the auxiliary refresh; a
sorbing tracer's total-sorbed constraint against the inverse isotherms.
Tolerances as the other parity tests: free ions and sorbed totals 1e-10
relative, Jacobian entries 1e-12 of the terms summed into them."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN
from gpu_cases import host_arrays, jacobian_bound, speciated_state
from pflotran_amd import abi, runtime, synth
from pflotran_amd.database import KDRxn, build_tables, read_database

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.fixture(scope="module")
def kdnet():
    return synth.with_isotherms(synth.random_network(15, 50, 3, ionx=True))


def test_isotherm_residual_jacobian_parity(kdnet):
    t = kdnet
    n = 300
    cells, _, _, ok = speciated_state(t, n)
    assert ok.all() and np.all(np.abs(cells["total_sorb"][:, t.kd_specid]) > 0.0)
    cells["pri_molal"][:] *= np.exp(0.3 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    o = synth.react_opts()
    res_o, jac_o = orc.residual_jacobian(t, cells, o)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    res_g, jac_g = runtime.residual_jacobian(tab, dc, synth.react_opts())
    assert np.all(np.abs(jac_g - jac_o) <= jacobian_bound(t, cells, jac_o))
    psv = 0.25 * 1.0 * 1000.0 * 1.0
    acc = (psv * np.abs(cells["total"]) + np.abs(cells["total_sorb"])) / synth.STEP_DT
    assert np.max(np.abs(res_g - res_o) / (acc + np.abs(res_o) + 1e-300)) <= 1e-12


@pytest.mark.parametrize("guess", [abi.PFX_GUESS_DEFAULT, abi.PFX_GUESS_PREVIOUS])
def test_isotherm_reaction_step_parity(kdnet, guess):
    t = kdnet
    n = 400
    cells, _, _, _ = speciated_state(t, n)
    # old sorbed amounts away from the new state's: half of them dropped for
    # the isotherm species that are not exchange cations (the exchanger's
    # sorbed charge is fixed at its CEC, so halving it has no solution)
    cats = set(t.ionx_cationid[0, 1:1 + t.ionx_cationid[0, 0]].tolist())
    cols = [j for j in t.kd_specid if j not in cats]
    assert cols
    cells["total_sorb"][::2, cols] *= 0.5
    ref = cells.copy()
    rc_o, _ = orc.react_batch(t, ref, synth.react_opts(guess_mode=guess))
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    rc_g, st = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=guess))
    out = dc.to_host()
    assert np.array_equal(out["status"], ref["status"]) and (ref["status"] == 0).mean() > 0.9, st.as_dict()
    ok = ref["status"] == 0
    assert rel_err(out["pri_molal"][ok], ref["pri_molal"][ok]) <= 1e-10
    assert rel_err(out["total_sorb"][ok], ref["total_sorb"][ok]) <= 1e-10
    assert rel_err(out["total"][ok], ref["total"][ok]) <= 1e-10


def test_isotherm_auxvar_parity(kdnet):
    t = kdnet
    n = 257
    cells, _, _, _ = speciated_state(t, n)
    cells["pri_molal"][:] *= np.exp(0.2 * np.cos(np.arange(n * t.N)).reshape(n, t.N))
    ref = cells.copy()
    orc.auxvar(t, ref, synth.react_opts())
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.auxvar(tab, dc, 0, synth.react_opts())
    out = dc.to_host()
    assert rel_err(out["total_sorb"], ref["total_sorb"]) <= 1e-12
    assert rel_err(out["total"], ref["total"]) <= 1e-12


@pytest.mark.parametrize("kind", [abi.PFX_SORPTION_LINEAR, abi.PFX_SORPTION_LANGMUIR, abi.PFX_SORPTION_FREUNDLICH])
def test_total_sorbed_constraint_on_the_device(kind):
    kd, b, nf = {abi.PFX_SORPTION_LINEAR: (40.0, 0.0, 0.0), abi.PFX_SORPTION_LANGMUIR: (3e3, 0.2, 0.0),
                 abi.PFX_SORPTION_FREUNDLICH: (5.0, 0.0, 1.7)}[kind]
    db = read_database(os.path.join(GOLDEN, "db", "hanford_subset.dat"))
    t = build_tables(db, ["A(aq)"], isotherms=[KDRxn("A(aq)", kd, kind, b, nf)], log_formulation=False)
    target = np.array([1e-6, 1e-4, 1e-2 if kind != abi.PFX_SORPTION_LANGMUIR else 0.1])
    n = len(target)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(synth.host_state_arrays(t, n, False), n)
    rc, st = runtime.equilibrate(tab, dc, [abi.PFX_CONSTRAINT_TOTAL_SORB], target.reshape(-1, 1),
                                 synth.react_opts(porosity=0.25))
    assert rc == 0, st.as_dict()
    out = dc.to_host()
    want = {abi.PFX_SORPTION_LINEAR: target / kd, abi.PFX_SORPTION_LANGMUIR: target / (kd * (b - target)),
            abi.PFX_SORPTION_FREUNDLICH: (target / kd) ** nf}[kind]
    assert np.all(np.abs(out["pri_molal"][:, 0] / want - 1.0) <= 1e-10)
    assert np.all(np.abs(out["total_sorb"][:, 0] / target - 1.0) <= 1e-12)
