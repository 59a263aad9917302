"""The oracle's kinetic mineral rate (RKineticMineral, reaction_mineral.F90:
564-1000) in every option branch the reference carries: Arrhenius
temperature dependence, affinity threshold, rate limiter, affinity power and
Temkin constant (MINERAL_KINETICS keywords, reaction_mineral.F90:151-210).

Pins:
- the rate against the transition-state law written out from the
  documentation (docs/user_manual/governing_eqns.tex:916-961) with the
  reference's conventions: Q/K capped at 1e3 above ln(Q/K) = 6.90776, the
  affinity factor 1 - (Q/K)^(1/sigma), precipitation suppressed below the
  threshold, the limiter's a / (1 + (1 - a) / limiter), |a|^beta, the
  Arrhenius factor exp(Ea/R (1/298.15 - 1/T)), zero rate for a mineral that
  is absent and undersaturated;
- the mineral part of the analytic Jacobian against central differences of
  the residual, where the reference's expression is the derivative (the
  plain law, Arrhenius, threshold, Temkin, and affinity power when
  undersaturated), up to the reference's units factor: it multiplies
  dQ/dm by the water density / 1000 (reaction_mineral.F90:862-866, "L water
  to kg water"), though Q is formed from molal activities, so its mineral
  Jacobian is the derivative times rho_w/1000 (0.99716 at 25 C);
- the two places where it is not, so that a "fix" would show: with
  AFFINITY_POWER the reference's dI/d(Q/K) = -I beta/|a| (:848) has the
  opposite sign of the derivative when the mineral is supersaturated
  (a < 0), and the limiter's factor (1 + (Q/K)/lim/den)/den with den = 1 +
  (1 - a_lim)/lim (:882-895) differs from the derivative of the limited
  affinity."""
import os

import numpy as np
import pytest

import oracle as orc
from chemdefs import DB
from pflotran_amd import abi
from pflotran_amd.database import MineralKinetics, build_tables, read_database

LOG_TO_LN = 2.30258509299
R_GAS = 8.31446
DEN = 997.16
PRIM = ["H+", "HCO3-", "Ca++"]
CPLX = ["OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"]


def tables(**kw):
    db = read_database(os.path.join(DB, "calcite.dat"))
    return build_tables(db, PRIM, CPLX, {"Calcite": MineralKinetics(1e-9, **kw)}, act_freq=abi.PFX_ACT_OFF)


def cells_at(t, log_omega, temp=40.0, vf=0.1, area=7.0):
    """Cells whose calcite saturation index is log_omega (Ca++ set from it at
    unit activity; H+ and HCO3- fixed)."""
    n = len(log_omega)
    c = orc.HostCells.allocate(t, n, temperature=True)
    k = 0
    ids = [int(x) for x in t.mnrl_specid[k, 1:1 + t.mnrl_specid[k, 0]]]
    nu = t.mnrl_stoich[k, :len(ids)]
    m = np.tile([10 ** -7.3, 2e-3, 1e-3], (n, 1))
    ca = PRIM.index("Ca++")
    rest = sum(nu[i] * np.log10(m[:, ids[i]]) for i in range(len(ids)) if ids[i] != ca)
    nu_ca = nu[ids.index(ca)]
    m[:, ca] = 10 ** ((np.asarray(log_omega) + t.mnrl_logK[k] - rest) / nu_ca)
    c["pri_molal"][:] = m
    c["temperature"][:] = temp
    c["mnrl_volfrac"][:, 0] = vf
    c["mnrl_area"][:, 0] = area
    return c


def tst_rate(t, c, k25, Ea=0.0, thr=0.0, lim=0.0, beta=None, sigma=None):
    """The transition-state rate law (mol/m^3 bulk/s) at the cells' state."""
    k = 0
    ids = [int(x) for x in t.mnrl_specid[k, 1:1 + t.mnrl_specid[k, 0]]]
    nu = t.mnrl_stoich[k, :len(ids)]
    lnq = -t.mnrl_logK[k] * LOG_TO_LN + sum(nu[i] * np.log(c["pri_molal"][:, ids[i]] * c["pri_act_coef"][:, ids[i]])
                                             for i in range(len(ids)))
    qk = np.where(lnq <= 6.90776, np.exp(np.minimum(lnq, 6.90776)), 1e3)
    a = 1.0 - qk ** (1.0 / (sigma if sigma is not None else 1.0))
    sgn = np.where(a >= 0.0, 1.0, -1.0)
    if lim > 0.0:
        a = a / (1.0 + (1.0 - a) / lim)
    kT = k25 * np.exp(Ea / R_GAS * (1.0 / 298.15 - 1.0 / (c["temperature"] + 273.15))) if Ea > 0 else k25
    rate = -c["mnrl_area"][:, 0] * sgn * np.abs(a) ** (beta if beta is not None else 1.0) * kT
    present = (c["mnrl_volfrac"][:, 0] > 0.0) | (sgn < 0.0)
    below = (thr > 0.0) & (sgn < 0.0) & (qk < thr)
    return np.where(present & ~below, rate, 0.0), qk, a


def oracle_rate_and_jac(t, c, o):
    cc = c.copy()
    r, J = orc.residual_jacobian(t, cc, o, fixed_accum=np.zeros_like(c["total"]))
    return cc["mnrl_rate"][:, 0].copy(), r, J


CASES = [dict(), dict(Ea=41870.0), dict(thr=3.0), dict(lim=0.7), dict(beta=1.7), dict(sigma=2.5),
         dict(beta=0.6, sigma=1.8, Ea=25000.0)]


@pytest.mark.parametrize("case", CASES, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()) or "plain")
def test_rate_matches_the_tst_law(case):
    kw = dict(activation_energy=case.get("Ea", 0.0), affinity_threshold=case.get("thr", 0.0),
              rate_limiter=case.get("lim", 0.0), affinity_power=case.get("beta"), temkin=case.get("sigma"))
    t = tables(**kw)
    log_omega = np.array([-3.0, -1.0, -0.2, 0.1, 0.4, 0.6, 1.5, 2.5, 3.2])
    c = cells_at(t, log_omega)
    o = abi.default_opts(dt=100.0, den_kg=DEN)
    rate, _, _ = oracle_rate_and_jac(t, c, o)
    want, _, _ = tst_rate(t, c, 1e-9, **case)
    assert np.all(np.abs(rate - want) <= 1e-13 * np.maximum(np.abs(want), 1e-30)), (rate, want)
    # an absent mineral dissolves nothing
    c0 = cells_at(t, log_omega, vf=0.0)
    rate0, _, _ = oracle_rate_and_jac(t, c0, o)
    want0, _, _ = tst_rate(t, c0, 1e-9, **case)
    assert np.all(rate0[log_omega < 0] == 0.0) and np.allclose(rate0, want0, rtol=1e-13, atol=0.0)
    if case.get("thr"):
        # supersaturated below the threshold: no precipitation
        assert np.all(rate[(log_omega > 0) & (10 ** log_omega < case["thr"])] == 0.0)
        assert np.all(rate[10 ** log_omega > case["thr"]] > 0.0)   # precipitation


def mineral_jacobians(case, log_omega):
    """(analytic, central-difference) mineral part of the Jacobian (J with
    the mineral less J with a zero rate constant), and the state's Q/K and
    limited affinity."""
    kw = dict(activation_energy=case.get("Ea", 0.0), affinity_threshold=case.get("thr", 0.0),
              rate_limiter=case.get("lim", 0.0), affinity_power=case.get("beta"), temkin=case.get("sigma"))
    t = tables(**kw)
    t0 = tables(**kw)
    t0.mnrl_rate_constant[:] = 0.0
    c = cells_at(t, log_omega)
    o = abi.default_opts(dt=100.0, den_kg=DEN)
    _, _, J = oracle_rate_and_jac(t, c, o)
    _, _, J0 = oracle_rate_and_jac(t0, c, o)
    Jfd = np.zeros_like(J)
    for j in range(t.N):
        h = 1e-4 * c["pri_molal"][:, j]   # central differences: h^2 truncation, rounding of the accumulation below 1e-5
        cp, cm = c.copy(), c.copy()
        cp["pri_molal"][:, j] += h
        cm["pri_molal"][:, j] -= h
        _, rp, _ = oracle_rate_and_jac(t, cp, o)
        _, rm, _ = oracle_rate_and_jac(t, cm, o)
        _, rp0, _ = oracle_rate_and_jac(t0, cp, o)
        _, rm0, _ = oracle_rate_and_jac(t0, cm, o)
        Jfd[:, :, j] = ((rp - rp0) - (rm - rm0)) / (2 * h[:, None])
    _, qk, a = tst_rate(t, c, 1e-9, **case)
    return J - J0, Jfd, qk, a


def close(a, b, rel=1e-4):
    """a equals b times the reference's units factor rho_w/1000."""
    b = b * DEN * 1e-3
    scale = np.max(np.abs(b), axis=(1, 2), keepdims=True)
    return np.all(np.abs(a - b) <= rel * scale)


@pytest.mark.parametrize("case", [dict(), dict(Ea=41870.0), dict(thr=3.0), dict(sigma=2.5), dict(beta=1.7)],
                         ids=["plain", "arrhenius", "threshold", "temkin", "power"])
def test_mineral_jacobian_is_the_derivative(case):
    # undersaturated states, and supersaturated ones away from the threshold
    log_omega = np.array([-2.0, -0.5, -0.1]) if "beta" in case else np.array([-2.0, -0.5, -0.1, 0.7, 1.8, 2.6])
    Jm, Jfd, _, _ = mineral_jacobians(case, log_omega)
    assert np.max(np.abs(Jm)) > 0.0
    assert close(Jm, Jfd)
    assert not close(Jm, Jfd * 1000.0 / DEN)   # the density factor is there


def test_affinity_power_derivative_sign_when_supersaturated():
    # the reference's -I beta/|a| is -(dI/dQK) where a < 0
    Jm, Jfd, _, a = mineral_jacobians(dict(beta=1.7), np.array([0.3, 1.0, 2.2]))
    assert np.all(a < 0.0)
    assert close(Jm, -Jfd)


def test_rate_limiter_jacobian_is_the_reference_expression():
    lim = 0.7
    log_omega = np.array([-1.5, -0.3, 0.4, 1.6])
    Jm, Jfd, qk, a_lim = mineral_jacobians(dict(lim=lim), log_omega)
    # the exact derivative of a0/(1 + (1 - a0)/lim), a0 = 1 - Q/K, is
    # -(1 + 1/lim)/(1 + Q/K/lim)^2; the reference's factor replaces it by
    # -(1 + Q/K/lim/den)/den with den = 1 + (1 - a_lim)/lim
    den = 1.0 + (1.0 - a_lim) / lim
    ref = (1.0 + qk / lim / den) / den
    exact = (1.0 + 1.0 / lim) / (1.0 + qk / lim) ** 2
    ratio = (ref / exact)[:, None, None]
    assert not close(Jm, Jfd)
    assert close(Jm, Jfd * ratio)


# ---------------------------------------------------------------------------
# rate prefactors (PREFACTOR / PREFACTOR_SPECIES, reaction_mineral.F90:227-360,
# 743-786, 887-944) and MIN_SCALE_FACTOR (:704-716, 800-802, 848-858)
PREF = [dict(rate=3.5e-8, species=[("HCO3-", 0.31, 0.0, 0.0), ("H+", 0.37, 1.0, 1e9)]),
        dict(rate=None, Ea=30000.0, species=[("Ca++", 0.5, 2.0, 40.0)])]


def pref_tables(prefactors=PREF, msf=None, rate_constant=1e-9):
    from pflotran_amd.database import Prefactor, PrefactorSpecies
    pf = [Prefactor([PrefactorSpecies(n, a, b, k) for n, a, b, k in p["species"]], rate=p.get("rate"),
                    activation_energy=p.get("Ea")) for p in prefactors]
    db = read_database(os.path.join(DB, "calcite.dat"))
    return build_tables(db, PRIM, CPLX, {"Calcite": MineralKinetics(rate_constant, prefactors=pf,
                                                                    min_scale_factor=msf)},
                        act_freq=abi.PFX_ACT_OFF)


def pref_law(c, prefactors=PREF, k_outer=1e-9):
    """sum_p k_p A_p(T) prod_s a_s^alpha / (1 + atten a_s^beta), the
    prefactor rate constant written out (activities = molalities here)."""
    total = np.zeros(c.ncells)
    one = np.ones(c.ncells)
    for p in prefactors:
        f = one.copy()
        for name, alpha, beta, atten in p["species"]:
            a = c["pri_molal"][:, PRIM.index(name)] * c["pri_act_coef"][:, PRIM.index(name)]
            f = f * a ** alpha / (1.0 + atten * a ** beta)
        kp = k_outer if p.get("rate") is None else p["rate"]
        Ea = p.get("Ea", 0.0) or 0.0
        arr = np.exp(Ea / R_GAS * (1.0 / 298.15 - 1.0 / (c["temperature"] + 273.15))) if Ea > 0 else 1.0
        total = total + kp * arr * f
    return total


def test_prefactor_rate_matches_the_law():
    t = pref_tables()
    assert t.mnrl_num_prefactors.tolist() == [2] and t.mnrl_pref_rate[0, 1] == 1e-9   # outer rate inherited
    log_omega = np.array([-3.0, -1.0, -0.2, 0.4, 1.5, 3.2])
    c = cells_at(t, log_omega)
    o = abi.default_opts(dt=100.0, den_kg=DEN)
    rate, _, _ = oracle_rate_and_jac(t, c, o)
    _, qk, a = tst_rate(t, c, 1.0)
    want = -c["mnrl_area"][:, 0] * np.sign(a) * np.abs(a) * pref_law(c)
    assert np.all(np.abs(rate - want) <= 1e-13 * np.abs(want)), (rate, want)
    # without the attenuation term the H+ factor is a^alpha alone
    t2 = pref_tables([dict(rate=2e-8, species=[("H+", 0.37, 1.0, 0.0)])])
    rate2, _, _ = oracle_rate_and_jac(t2, c, o)
    want2 = -c["mnrl_area"][:, 0] * np.sign(a) * np.abs(a) * 2e-8 * (c["pri_molal"][:, 0] ** 0.37)
    assert np.all(np.abs(rate2 - want2) <= 1e-13 * np.abs(want2))


def test_min_scale_factor_rate_matches_the_law():
    f = 2.5
    t = pref_tables([], msf=f)
    log_omega = np.array([-2.0, -0.4, 0.3, 1.1])
    c = cells_at(t, log_omega)
    rate, _, _ = oracle_rate_and_jac(t, c, abi.default_opts(dt=100.0, den_kg=DEN))
    _, qk, _ = tst_rate(t, c, 1.0)
    a = 1.0 - qk ** (1.0 / f)
    want = -c["mnrl_area"][:, 0] / f * np.sign(a) * np.abs(a) * 1e-9
    assert np.all(np.abs(rate - want) <= 1e-13 * np.abs(want))
    # with a Temkin constant the exponent is 1 / (f sigma)
    from pflotran_amd.database import build_tables as bt
    db = read_database(os.path.join(DB, "calcite.dat"))
    t3 = bt(db, PRIM, CPLX, {"Calcite": MineralKinetics(1e-9, temkin=1.5, min_scale_factor=f)},
            act_freq=abi.PFX_ACT_OFF)
    rate3, _, _ = oracle_rate_and_jac(t3, c, abi.default_opts(dt=100.0, den_kg=DEN))
    a3 = 1.0 - qk ** (1.0 / (f * 1.5))
    want3 = -c["mnrl_area"][:, 0] / f * np.sign(a3) * np.abs(a3) * 1e-9
    assert np.all(np.abs(rate3 - want3) <= 1e-13 * np.abs(want3))


@pytest.mark.parametrize("which", ["prefactor", "min_scale_factor"])
def test_prefactor_and_scale_factor_jacobian_is_the_derivative(which):
    # at a water density of 1000 kg/m^3 the reference's rho/1000 factor on
    # the Q/K part is 1, and the prefactor part has none: the analytic
    # mineral Jacobian is the derivative of the residual
    log_omega = np.array([-2.0, -0.5, -0.1, 0.7, 1.8])
    # the mineral's part: J less J with no mineral surface (a zero rate
    # constant would make the reference's Im / sum 0/0)
    t = pref_tables() if which == "prefactor" else pref_tables([], msf=2.5)
    c = cells_at(t, log_omega)
    # a long step: the accumulation's rounding over the H+ column's tiny
    # difference step stays below the mineral terms
    o = abi.default_opts(dt=1e7, den_kg=1000.0)

    def bare(x):
        x = x.copy()
        x["mnrl_area"][:] = 0.0
        return x
    _, _, J = oracle_rate_and_jac(t, c, o)
    _, _, J0 = oracle_rate_and_jac(t, bare(c), o)
    Jm = J - J0
    Jfd = np.zeros_like(J)
    for j in range(t.N):
        h = 1e-4 * c["pri_molal"][:, j]
        cp, cm = c.copy(), c.copy()
        cp["pri_molal"][:, j] += h
        cm["pri_molal"][:, j] -= h
        _, rp, _ = oracle_rate_and_jac(t, cp, o)
        _, rm, _ = oracle_rate_and_jac(t, cm, o)
        _, rp0, _ = oracle_rate_and_jac(t, bare(cp), o)
        _, rm0, _ = oracle_rate_and_jac(t, bare(cm), o)
        Jfd[:, :, j] = ((rp - rp0) - (rm - rm0)) / (2 * h[:, None])
    scale = np.max(np.abs(Jfd), axis=(1, 2), keepdims=True)
    assert np.max(np.abs(Jm)) > 0.0
    assert np.all(np.abs(Jm - Jfd) <= 1e-4 * scale)
    if which == "prefactor":
        # the H+ column carries the prefactor's derivative: dropping it breaks
        # the match (a plausible omission)
        Jq = Jm.copy()
        Jq[:, :, 0] = Jfd[:, :, 0] * 0.0
        assert not np.all(np.abs(Jq - Jfd) <= 1e-4 * scale)
