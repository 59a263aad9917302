"""HIP kernels (through the C-ABI) against the CPU oracle on the same seeded
inputs.  Tolerances: free ions 1e-10 relative, Jacobian entries 1e-12
relative to the largest entry of their row, residual 1e-12 relative to the
accumulation magnitude."""
import numpy as np
import pytest

import oracle as orc
from chemdefs import calcite_kinetics, ca_carbonate, carbonate, carbonate_column, valocchi
from conftest import gold
from gpu_cases import host_arrays, jacobian_bound, speciated_state
from pflotran_amd import abi, runtime, synth
from pflotran_amd.database import water_density_ifc67

pytestmark = pytest.mark.gpu

DEN = water_density_ifc67()


def rel_err(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.fixture(scope="module")
def net15():
    return synth.random_network(15, 50, 3, ionx=True)


def test_residual_jacobian_parity(net15):
    t = net15
    n = 333
    cells, _, _, ok = speciated_state(t, n)
    assert ok.all()
    o = synth.react_opts()
    # evaluate at a perturbed iterate so kinetics and exchange are far from balance
    cells["pri_molal"][:] *= np.exp(0.3 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    res_o, jac_o = orc.residual_jacobian(t, cells, o)  # also fills sec_molal at the state
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    res_g, jac_g = runtime.residual_jacobian(tab, dc, synth.react_opts())
    assert np.all(np.abs(jac_g - jac_o) <= jacobian_bound(t, cells, jac_o))
    psv = 0.25 * 1.0 * 1000.0 * 1.0
    acc = psv * np.abs(cells["total"]) / synth.STEP_DT
    assert np.max(np.abs(res_g - res_o) / (acc + np.abs(res_o) + 1e-300)) <= 1e-12


def test_residual_jacobian_parity_numerical():
    # NUMERICAL_JACOBIAN: the reaction block by forward differences
    # (reaction.F90:3625-3684).  Each implementation's rounding of the
    # residual sums, a few ulp of S_i = sum_k |nu_ik Im_k| V, is divided by the
    # perturbation 1e-5 m_j: entries are held to 1e-10 S_i / m_j (about 4 ulp
    # each side) plus 1e-8 of the differenced column
    t = synth.random_network(15, 50, 3, ionx=True)
    t.numerical_jacobian = 1
    n = 200
    cells, _, _, ok = speciated_state(t, n)
    cells["pri_molal"][:] *= np.exp(0.3 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    o = synth.react_opts()
    res_o, jac_o = orc.residual_jacobian(t, cells, o)
    bare = cells.copy()
    bare["mnrl_area"][:] = 0.0
    _, jac_0 = orc.residual_jacobian(t, bare, o)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    res_g, jac_g = runtime.residual_jacobian(tab, dc, synth.react_opts())
    fdscale = np.max(np.abs(jac_o - jac_0), axis=1, keepdims=True)
    assert np.count_nonzero(fdscale) > 0
    _, _ = orc.residual_jacobian(t, cells, o)  # sec_molal and mineral rates at the state
    nu = np.zeros((t.K, t.N))
    for k in range(t.K):
        for s_ in range(t.mnrl_specid[k, 0]):
            nu[k, t.mnrl_specid[k, 1 + s_]] = t.mnrl_stoich[k, s_]
    S = np.abs(cells["mnrl_rate"]) @ np.abs(nu)               # [cell][row], V = 1 m^3
    noise = 1e-10 * S[:, :, None] / cells["pri_molal"][:, None, :]
    excess = np.abs(jac_g - jac_o) - jacobian_bound(t, cells, jac_o) - noise - 1e-8 * fdscale
    worst = np.unravel_index(np.argmax(excess), excess.shape)
    assert excess[worst] <= 0.0, (worst, jac_g[worst], jac_o[worst], fdscale[worst[0], 0, worst[2]])
    psv = 0.25 * 1.0 * 1000.0 * 1.0
    acc = psv * np.abs(cells["total"]) / synth.STEP_DT
    assert np.max(np.abs(res_g - res_o) / (acc + np.abs(res_o) + 1e-300)) <= 1e-12


def test_react_parity_default_guess(net15):
    t = net15
    n = 512
    cells, _, _, ok = speciated_state(t, n)
    assert ok.all()
    o = synth.react_opts()
    ref = cells.copy()
    rc_o, st_o = orc.react_batch(t, ref, o)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    rc_g, st_g = runtime.react_batch(tab, dc, synth.react_opts())
    out = dc.to_host()
    assert rc_g == rc_o == 0, (st_g.as_dict(), st_o.as_dict())
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10
    assert rel_err(out["total"], ref["total"]) <= 1e-10
    assert rel_err(out["pri_act_coef"], ref["pri_act_coef"]) <= 1e-12
    assert rel_err(out["total_sorb"], ref["total_sorb"]) <= 1e-10
    assert np.max(np.abs(out["mnrl_rate"] - ref["mnrl_rate"]) / np.maximum(np.abs(ref["mnrl_rate"]), 1e-30)) <= 1e-9
    assert np.all(out["status"] == 0)
    # iteration counts agree closely (rounding can move a cell across the tolerance)
    assert np.mean(np.abs(out["num_iterations"] - ref["num_iterations"])) < 0.5


def test_react_parity_capped_mineral_derivative(net15):
    # PFX_JAC_CAPPED_QK: the oracle's Newton with the same option, and the
    # same solution as the reference Jacobian's in fewer iterations
    t = net15
    n = 512
    cells, _, _, _ = speciated_state(t, n)
    oc = dict(jacobian_flags=abi.PFX_JAC_CAPPED_QK)
    ref = cells.copy()
    orc.react_batch(t, ref, synth.react_opts(**oc))
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.react_batch(tab, dc, synth.react_opts(**oc))
    out = dc.to_host()
    assert np.array_equal(out["status"], ref["status"]) and np.all(out["status"] == 0)
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10
    assert np.mean(np.abs(out["num_iterations"] - ref["num_iterations"])) < 0.5
    dr = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.react_batch(tab, dr, synth.react_opts())
    plain = dr.to_host()
    assert rel_err(out["pri_molal"], plain["pri_molal"]) <= 1e-9
    assert out["num_iterations"].mean() < 0.8 * plain["num_iterations"].mean()


@pytest.mark.parametrize("n", [1, 7, 33, 1000])
def test_react_ragged_and_previous_guess(net15, n):
    t = net15
    cells, _, _, ok = speciated_state(t, n, seed=7)
    o = synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS)
    ref = cells.copy()
    orc.react_batch(t, ref, o)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    out = dc.to_host()
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10


@pytest.mark.parametrize("alg", [abi.PFX_ACT_ALG_LAG, abi.PFX_ACT_ALG_NEWTON])
def test_equilibrate_parity(net15, alg):
    t = net15
    if alg != t.act_alg:
        t = synth.random_network(15, 50, 3, ionx=True)
        t.act_alg = alg
    n = 256
    cells, ctype, conc, ok = speciated_state(t, n, seed=11)
    tab = runtime.Tables(t)
    arr = host_arrays(cells)
    arr["pri_molal"][:] = 1e-9
    dc = runtime.DeviceCells.from_host(arr, n)
    o = synth.react_opts()
    rc, st = runtime.equilibrate(tab, dc, ctype, conc, o)
    assert rc == 0, st.as_dict()
    # the initial-condition refresh: unit coefficients, two activity passes
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    dc.t["ionx_ref_sorbed"].fill_(1e-9)
    runtime.auxvar(tab, dc, 2, synth.react_opts())
    out = dc.to_host()
    assert rel_err(out["pri_molal"], cells["pri_molal"]) <= 1e-10
    assert rel_err(out["pri_act_coef"], cells["pri_act_coef"]) <= 1e-10
    assert rel_err(out["total"][:, 1:], cells["total"][:, 1:]) <= 1e-10


@pytest.mark.parametrize("freq", [abi.PFX_ACT_TIMESTEP, abi.PFX_ACT_NEWTON_ITER])
@pytest.mark.parametrize("alg", [abi.PFX_ACT_ALG_LAG, abi.PFX_ACT_ALG_NEWTON])
def test_react_parity_activity_algorithms(alg, freq):
    # activity coefficients once per step or every Newton iteration, by the
    # one-pass LAG form or the ionic-strength iteration (reaction.F90:3812-4053).
    # Updates every iteration start from the previous free ions, as RReact
    # does: from the 1e-9 default most cells stop in the exchange or
    # ionic-strength iterations, in the oracle and on the GPU alike.
    t = synth.random_network(15, 50, 3, ionx=True, act_freq=freq)
    t.act_alg = alg
    n = 256
    cells, _, _, _ = speciated_state(t, n, seed=23)
    guess = abi.PFX_GUESS_PREVIOUS if freq == abi.PFX_ACT_NEWTON_ITER else abi.PFX_GUESS_DEFAULT
    ref = cells.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=guess))
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.react_batch(tab, dc, synth.react_opts(guess_mode=guess))
    out = dc.to_host()
    assert np.array_equal(out["status"], ref["status"])
    ok = ref["status"] == 0
    assert ok.mean() > 0.9
    keys = ("pri_molal", "pri_act_coef", "sec_act_coef")
    bound = {k: np.full(ok.sum(), 1e-10) for k in keys}
    if alg == abi.PFX_ACT_ALG_NEWTON and freq == abi.PFX_ACT_NEWTON_ITER:
        # The ionic-strength iteration stops at |I - I'| < 1e-6 I
        # (reaction.F90:3884), so the coefficients are a step function of
        # m, re-evaluated at every Newton iterate: where the inner count
        # flips, a rounding-level change of the input moves the solution
        # beyond 1e-10.  Per cell, the oracle's own spread over two inputs
        # perturbed by 1e-15 relative bounds the GPU's difference (x4).
        for s in (1, 2):
            p = cells.copy()
            p["pri_molal"][:] *= 1.0 + 1e-15 * np.random.default_rng(s).uniform(-1, 1, p["pri_molal"].shape)
            orc.react_batch(t, p, synth.react_opts(guess_mode=guess))
            for k in keys:
                spread = np.max(np.abs(p[k][ok] - ref[k][ok]) / np.abs(ref[k][ok]), axis=1)
                bound[k] = np.maximum(bound[k], 4.0 * spread)
    for k in keys:
        err = np.max(np.abs(out[k][ok] - ref[k][ok]) / np.abs(ref[k][ok]), axis=1)
        assert np.all(err <= bound[k]), (k, float(err.max()), int(np.argmax(err / bound[k])))
        assert np.mean(bound[k] > 1e-10) < 0.1   # the perturbation-sensitive cells are few
    assert np.mean(np.abs(out["num_iterations"][ok] - ref["num_iterations"][ok])) < 0.5 or \
        (alg == abi.PFX_ACT_ALG_NEWTON and freq == abi.PFX_ACT_NEWTON_ITER)


def _gpu_speciate(t, setup, ncells=1):
    tab = runtime.Tables(t)
    arr = synth.host_state_arrays(t, ncells, with_temperature=False)
    if "vf" in setup:
        arr["mnrl_volfrac"][:, 0] = setup["vf"]
        arr["mnrl_area"][:, 0] = setup["area"]
    dc = runtime.DeviceCells.from_host(arr, ncells)
    o = synth.react_opts(porosity=setup["porosity"])
    conc = np.tile(np.array(setup["conc"], dtype=np.float64), (ncells, 1))
    rc, st = runtime.equilibrate(tab, dc, setup["ctype"], conc, o, rxn=setup.get("rxn"))
    assert rc == 0, st.as_dict()
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    if dc.t["ionx_ref_sorbed"] is not None:
        dc.t["ionx_ref_sorbed"].fill_(1e-9)
    passes = 2 if t.act_freq != abi.PFX_ACT_OFF else 0
    runtime.auxvar(tab, dc, passes, synth.react_opts(porosity=setup["porosity"]))
    return tab, dc


@pytest.mark.parametrize("act", [False, True])
def test_gpu_golds_speciation(golds, act):
    for name, (t, setup), ph_col, species in (
            ("carbonate-debye-huckel-activity" if act else "carbonate-unit-activity", carbonate(act), 0,
             ["H+", "HCO3-"]),
            ("ca-carbonate-debye-huckel-activity" if act else "ca-carbonate-unit-activity", ca_carbonate(act), 2,
             ["HCO3-", "Ca++", "H+"])):
        _, dc = _gpu_speciate(t, setup)
        out = dc.to_host()
        m, g = out["pri_molal"][0], out["pri_act_coef"][0]
        assert abs(-np.log10(g[ph_col] * m[ph_col]) - gold(golds, name, "GENERIC: pH")) < 1e-12
        for j, sp in enumerate(species):
            assert abs(out["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < 1e-12
            if act:
                assert abs(g[j] - gold(golds, name, f"GENERIC: Gamma {sp}")) < 1e-12


def test_gpu_gold_calcite_kinetics(golds):
    """calcite-kinetics: 500 reaction steps of 0.5 s on the GPU, each with the
    fixed accumulation refreshed from the current state."""
    name = "calcite-kinetics"
    t, setup = calcite_kinetics()
    tab, dc = _gpu_speciate(t, setup)
    o = synth.react_opts(porosity=0.5, guess_mode=abi.PFX_GUESS_PREVIOUS, rel_tol=1e-14, extra_iterations=1,
                         max_iterations=60, dt=0.5)
    for _ in range(500):
        runtime.auxvar(tab, dc, 0, synth.react_opts(porosity=0.5))  # accumulation of the old state
        rc, st = runtime.react_batch(tab, dc, synth.react_opts(porosity=0.5, guess_mode=abi.PFX_GUESS_PREVIOUS,
                                                               rel_tol=1e-14, extra_iterations=1,
                                                               max_iterations=60, dt=0.5))
        assert rc == 0
        runtime.update_kinetic_state(tab, dc, synth.react_opts(porosity=0.5, dt=0.5))
    runtime.auxvar(tab, dc, 0, synth.react_opts(porosity=0.5))
    out = dc.to_host()
    m, g = out["pri_molal"][0], out["pri_act_coef"][0]
    assert abs(-np.log10(g[0] * m[0]) - gold(golds, name, "GENERIC: pH")) < 1e-12
    for j, sp in enumerate(["H+", "HCO3-", "Ca++"]):
        assert abs(out["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < 1e-12
    assert abs(out["mnrl_volfrac"][0, 0] - gold(golds, name, "VOLUME_FRACTION: Calcite VF")) < 1e-12
    assert abs(out["mnrl_rate"][0, 0] - gold(golds, name, "RATE: Calcite Rate")) < 1e-12


def test_update_kinetic_state_parity(net15):
    # RUpdateKineticState after a converged step: the GPU update from the
    # step's returned rates against the oracle, which recomputes the rates
    # at the returned state
    t = net15
    n = 200
    cells, _, _, _ = speciated_state(t, n, seed=5)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    runtime.react_batch(tab, dc, synth.react_opts())
    state = orc.HostCells({k: (None if v is None else v.copy()) for k, v in dc.to_host().items()}, n)
    runtime.update_kinetic_state(tab, dc, synth.react_opts())
    orc.update_kinetic_state(t, state, synth.react_opts())
    vf0 = cells["mnrl_volfrac"]
    d_g = dc.t["mnrl_volfrac"].cpu().numpy() - vf0
    d_o = state["mnrl_volfrac"] - vf0
    assert np.count_nonzero(d_o) > 0
    # the change itself (rate * molar volume * dt) to 1e-9 of its size, plus
    # the rounding of vf0 + change
    assert np.all(np.abs(d_g - d_o) <= 1e-9 * np.abs(d_o) + 4e-16 * vf0)


def test_gpu_gold_valocchi_charge_balance(golds):
    """ion-exchange-valocchi: Na+/Ca++/Mg++ totals with Cl- by charge balance
    (constraint speciation on the GPU), then the sorbed totals of the
    exchanger at the solution."""
    from chemdefs import valocchi
    name = "ion-exchange-valocchi"
    t, setup = valocchi()
    tab, dc = _gpu_speciate(t, setup)
    out = dc.to_host()
    for j, sp in enumerate(["Na+", "Ca++", "Mg++", "Cl-"]):
        assert abs(out["total"][0][j] - gold(golds, name, f"CONCENTRATION: Total {sp}")) < 1e-12
        g_s = gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}")
        assert abs(out["total_sorb"][0][j] - g_s) <= 1e-12 * max(abs(g_s), 1.0)
    # charge balance holds at the solution: sum z_j T_j = 0
    assert abs(np.dot(t.Z, out["total"][0])) < 1e-11


@pytest.mark.parametrize("activity", [False, True])
def test_gpu_mineral_and_gas_constraints(activity):
    """The calcite decks' initial constraint (H+ free, HCO3- by CO2(g) at
    1e-3 bar, Ca++ by calcite) on the GPU against the oracle's solution."""
    from chemdefs import calcite_column_initial
    t, setup = calcite_column_initial(activity)
    tab, dc = _gpu_speciate(t, setup, ncells=3)
    out = dc.to_host()
    ref = orc.HostCells.allocate(t, 3)
    rc, _ = orc.equilibrate(t, setup["ctype"], np.tile(np.array(setup["conc"]), (3, 1)), ref,
                            synth.react_opts(porosity=setup["porosity"]), rxn=setup["rxn"])
    assert rc == 0
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10
    if not activity:
        la = np.log(out["pri_molal"])
        assert np.max(np.abs(-1.8487 * 2.30258509299 - la[:, 0] + la[:, 2] + la[:, 1])) <= 1e-10
        assert np.max(np.abs(7.8136 * 2.30258509299 + la[:, 0] + la[:, 1] - np.log(1e-3))) <= 1e-10


def test_gpu_sorbed_total_constraints(golds):
    """Ca++ and Mg++ constrained by the valocchi gold's sorbed totals on the
    GPU: the gold's state again, and the oracle's solution of the same rows."""
    from chemdefs import valocchi_sorbed
    name = "ion-exchange-valocchi"
    sorbed = [gold(golds, name, f"CONCENTRATION: Total Sorbed {sp}") for sp in ("Ca++", "Mg++")]
    t, setup = valocchi_sorbed(sorbed)
    tab, dc = _gpu_speciate(t, setup)
    out = dc.to_host()
    for j, sp in enumerate(["Na+", "Ca++", "Mg++", "Cl-"]):
        g_t = gold(golds, name, f"CONCENTRATION: Total {sp}")
        assert abs(out["total"][0][j] - g_t) <= 1e-9 * abs(g_t)
    ref = orc.HostCells.allocate(t, 1)
    ref["mnrl_volfrac"][0, 0] = setup["vf"]
    ref["mnrl_area"][0, 0] = setup["area"]
    rc, _ = orc.equilibrate(t, setup["ctype"], np.array(setup["conc"]), ref, synth.react_opts(porosity=setup["porosity"]))
    assert rc == 0
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10


def _with_mineral_options(t):
    """Every rate-law option of RKineticMineral on: mineral 0 with an
    activation energy, a precipitation threshold and a rate limiter,
    mineral 1 with an affinity power and a Temkin constant, mineral 2 with
    both defaults of the allocated arrays (1)."""
    t.mnrl_activation_energy[:] = [45000.0, 0.0, 20000.0]
    t.mnrl_affinity_threshold[:] = [2.0, 0.0, 0.0]
    t.mnrl_rate_limiter[:] = [0.5, 0.0, 0.0]
    t.mnrl_affinity_power = np.array([1.0, 1.5, 1.0])
    t.mnrl_temkin = np.array([1.0, 2.0, 1.0])
    return t


def test_mineral_option_parity():
    # the kernel's rate-law branches against the oracle's (which the
    # closed-form and finite-difference tests pin): residual, Jacobian and
    # rates at a perturbed iterate, then the reaction step's free ions
    t = _with_mineral_options(synth.random_network(15, 50, 3, ionx=True))
    n = 400
    cells, _, _, _ = speciated_state(t, n)
    cells["pri_molal"][:] *= np.exp(0.3 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    o = synth.react_opts()
    res_o, jac_o = orc.residual_jacobian(t, cells, o)
    rate_o = cells["mnrl_rate"].copy()
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    res_g, jac_g = runtime.residual_jacobian(tab, dc, synth.react_opts())
    rate_g = dc.to_host()["mnrl_rate"]
    assert np.count_nonzero(rate_o) > n and np.all(np.abs(rate_g - rate_o) <= 1e-12 * np.abs(rate_o))
    assert np.all(np.abs(jac_g - jac_o) <= jacobian_bound(t, cells, jac_o))
    psv = 0.25 * 1.0 * 1000.0 * 1.0
    acc = psv * np.abs(cells["total"]) / synth.STEP_DT
    assert np.max(np.abs(res_g - res_o) / (acc + np.abs(res_o) + 1e-300)) <= 1e-12
    # the reaction step from the previous free ions
    start, _, _, _ = speciated_state(t, n)
    ref = start.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    dr = runtime.DeviceCells.from_host(host_arrays(start), n)
    runtime.react_batch(tab, dr, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    out = dr.to_host()
    # with the reference's affinity-power derivative (of the wrong sign
    # where a mineral is supersaturated) and its limiter factor, Newton stalls
    # in some cells (about 3 in 10 here, the same cells in both)
    assert np.array_equal(out["status"], ref["status"]) and (ref["status"] == 0).mean() > 0.5
    ok = ref["status"] == 0
    assert rel_err(out["pri_molal"][ok], ref["pri_molal"][ok]) <= 1e-10


def _with_prefactors(t):
    """Rate prefactors and the min scale factor (reaction_mineral.F90:
    704-802): mineral 0 with two parallel prefactors (H+ with an
    attenuation, two other primaries), mineral 1 with one prefactor that
    has an activation energy and a scale factor 2, mineral 2 plain."""
    from pflotran_amd.database import MineralKinetics, Prefactor, PrefactorSpecies, prefactor_arrays
    names = t.primary_names
    kin = [MineralKinetics(float(t.mnrl_rate_constant[k]), activation_energy=float(t.mnrl_activation_energy[k]))
           for k in range(t.K)]
    kin[0].prefactors = [Prefactor([PrefactorSpecies(names[0], 0.37, 1.0, 1e5), PrefactorSpecies(names[3], 0.3)],
                                   rate=2.0 * t.mnrl_rate_constant[0]),
                         Prefactor([PrefactorSpecies(names[5], 0.5, 2.0, 40.0)], rate=t.mnrl_rate_constant[0])]
    kin[1].prefactors = [Prefactor([PrefactorSpecies(names[2], 0.25, 1.0, 10.0)], rate=t.mnrl_rate_constant[1],
                                   activation_energy=30000.0)]
    arr = prefactor_arrays(kin, {n: i for i, n in enumerate(names)})
    for k, v in arr.items():
        setattr(t, k, v)
    t.mnrl_min_scale_factor = np.array([1.0, 2.0, 1.0])
    return t


def test_prefactor_and_scale_factor_parity():
    t = _with_prefactors(synth.random_network(15, 50, 3, ionx=True))
    n = 400
    cells, _, _, _ = speciated_state(t, n)
    cells["pri_molal"][:] *= np.exp(0.3 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    o = synth.react_opts()
    res_o, jac_o = orc.residual_jacobian(t, cells, o)
    rate_o = cells["mnrl_rate"].copy()
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    res_g, jac_g = runtime.residual_jacobian(tab, dc, synth.react_opts())
    rate_g = dc.to_host()["mnrl_rate"]
    assert np.count_nonzero(rate_o[:, :2]) > n and np.all(np.abs(rate_g - rate_o) <= 1e-12 * np.abs(rate_o))
    assert np.all(np.abs(jac_g - jac_o) <= jacobian_bound(t, cells, jac_o))
    psv = 0.25 * 1.0 * 1000.0 * 1.0
    acc = psv * np.abs(cells["total"]) / synth.STEP_DT
    assert np.max(np.abs(res_g - res_o) / (acc + np.abs(res_o) + 1e-300)) <= 1e-12
    start, _, _, _ = speciated_state(t, n)
    ref = start.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    dr = runtime.DeviceCells.from_host(host_arrays(start), n)
    runtime.react_batch(tab, dr, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    out = dr.to_host()
    assert np.array_equal(out["status"], ref["status"]) and (ref["status"] == 0).mean() > 0.9
    ok = ref["status"] == 0
    assert rel_err(out["pri_molal"][ok], ref["pri_molal"][ok]) <= 1e-10


def test_equilibrate_from_a_free_ion_guess(net15):
    # FREE_ION_GUESS (reaction.F90:1476-1482): the constraint Newton starts
    # from cells->guess; the oracle from the same guess gives the same free
    # ions, which are the constraint's solution found from 1e-9 as well
    t = net15
    n = 256
    cells, ctype, conc, ok = speciated_state(t, n, seed=11)
    tab = runtime.Tables(t)
    arr = host_arrays(cells)
    arr["guess"][:] = cells["pri_molal"] * np.exp(0.5 * np.cos(np.arange(n * t.N)).reshape(n, t.N))
    arr["pri_molal"][:] = 1e-9
    o = synth.react_opts()
    o.guess_mode = abi.PFX_GUESS_ARRAY
    dc = runtime.DeviceCells.from_host(arr, n)
    rc, _ = runtime.equilibrate(tab, dc, ctype, conc, o)
    assert rc == 0
    out = dc.to_host()
    ref = orc.HostCells({k: (None if v is None else v.copy()) for k, v in arr.items()}, n)
    rc_o, its_o = orc.equilibrate(t, ctype, conc, ref, o)
    assert rc_o == 0
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-10
    assert np.mean(np.abs(out["num_iterations"] - its_o)) < 0.5
    d0 = runtime.DeviceCells.from_host(host_arrays(cells), n)
    d0.t["pri_molal"].fill_(1e-9)
    runtime.equilibrate(tab, d0, ctype, conc, synth.react_opts())
    plain = d0.to_host()
    # the constraint Newton stops on max|dm/m| < 1e-6 (or max|R| < 1e-12):
    # two starts end within that of each other
    assert rel_err(out["pri_molal"], plain["pri_molal"]) <= 1e-6
    assert out["num_iterations"].mean() < plain["num_iterations"].mean()
