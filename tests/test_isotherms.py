"""Isotherm (K_D) sorption: the deck reader's ISOTHERM_REACTIONS and the
oracle's RTotalSorbKD (reaction.F90:4220-4301) in its three types, pinned
by what fixes them independently of the oracle's own expressions:

- the reference's 1-cell decks state their K_D's meaning in their comments:
  default/batch/solute_KD_w_mineral gives 0.125786163522 mL/g scaled by the
  mineral's volume fraction 1, the solids 1 - 0.25 and 2650 kg/m^3 soil,
  which is the 250 kg water/m^3 bulk of solute_KD_wo_mineral, a
  retardation R = 1 + K_D / (phi s rho_w) = 2;
- the limits of the Langmuir isotherm (S -> b at saturation, K_D b m at
  low coverage), the Freundlich isotherm's homogeneity S(l m) = l^(1/n)
  S(m) and its n = 1 case, which is the linear isotherm;
- the derivatives against central differences of the residual;
- the reaction step of a sorbing tracer against the root of its mass
  balance found by bisection (phi s rho_w V m + S(m) V = the old amount);
- a total-sorbed constraint against the isotherm's inverse.

The column gold of default/column/solute_KD (111 steps) is reproduced by
the time loop tests with the other run golds."""
import os

import numpy as np
import pytest

import oracle as orc
from conftest import GOLDEN
from pflotran_amd import abi, synth
from pflotran_amd.database import KDRxn, build_tables, read_database
from pflotran_amd.deck import DeckError, parse_deck, read_deck

DB = os.path.join(GOLDEN, "db", "hanford_subset.dat")


def tracer(kind, kd, b=0.0, n=0.0, mineral=False):
    """A(aq) alone with one isotherm (and the inert mineral A(s) when the
    isotherm is scaled by it)."""
    from pflotran_amd.database import MineralKinetics
    db = read_database(DB)
    mins = {"A(s)": MineralKinetics(0.0)} if mineral else {}
    return build_tables(db, ["A(aq)"], [], mins,
                        isotherms=[KDRxn("A(aq)", kd, kind, b, n, "A(s)" if mineral else None)],
                        soil_particle_density=2650.0, log_formulation=False)


def sorbed(t, m, porosity=0.25, den=1000.0, vf=1.0):
    """(S, the oracle's dS/dm) at free-ion molalities m."""
    n = len(m)
    c = orc.HostCells.allocate(t, n, per_cell_scalars=True)
    c["pri_molal"][:, 0] = m
    c["porosity"][:] = porosity
    c["saturation"][:] = 1.0
    c["volume"][:] = 1.0
    c["den_kg"][:] = den
    if t.K:
        c["mnrl_volfrac"][:, 0] = vf
    o = abi.default_opts(dt=1.0, den_kg=den, porosity=porosity)
    orc.auxvar(t, c, o)
    S = c["total_sorb"][:, 0].copy()
    # d(residual)/dm = (phi s rho_w V + dS/dm V) / dt for the tracer, so the
    # Jacobian less its aqueous part is dS/dm (V = dt = 1)
    _, J = orc.residual_jacobian(t, c, o, fixed_accum=np.zeros((n, 1)))
    dS = J[:, 0, 0] - porosity * den
    return S, dS


M = np.logspace(-9, -1, 17)


# ---------------------------------------------------------------- deck


def test_deck_isotherms_and_units():
    d = read_deck(os.path.join(GOLDEN, "decks", "solute_KD.in"))
    (k,) = d.chemistry.isotherms
    assert (k.species, k.type, k.kd, k.mineral) == ("A(aq)", abi.PFX_SORPTION_LINEAR, 250.0, None)
    assert d.water_density() == 1000.0          # REFERENCE_DENSITY 1.d6 g/m^3
    w = read_deck(os.path.join(GOLDEN, "decks", "solute_KD_w_mineral.in"))
    (k,) = w.chemistry.isotherms
    assert k.mineral == "A(s)" and abs(k.kd - 0.125786163522) <= 1e-15   # mL/g is L/kg
    assert w.soil_particle_density() == 2650.0 and w.water_density() == 1000.0
    t = w.tables(w.database(DB))
    assert t.D == 1 and t.kd_mineral[0] == 0 and t.has_sorption


def _deck(isotherm_lines):
    return parse_deck("CHEMISTRY\n PRIMARY_SPECIES\n  A(aq)\n  B(aq)\n /\n SORPTION\n  ISOTHERM_REACTIONS\n"
                      + isotherm_lines + "  /\n /\nEND\n")


def test_deck_type_is_set_by_the_last_keyword_line():
    # every keyword line resets the type to LINEAR before it is read
    # (reaction.F90:539-541), so TYPE before DISTRIBUTION_COEFFICIENT is lost
    d = _deck("   A(aq)\n    TYPE LANGMUIR\n    DISTRIBUTION_COEFFICIENT 2.\n    /\n")
    assert d.chemistry.isotherms[0].type == abi.PFX_SORPTION_LINEAR
    d = _deck("   A(aq)\n    DISTRIBUTION_COEFFICIENT 2.\n    LANGMUIR_B 3.\n    /\n")
    assert d.chemistry.isotherms[0].type == abi.PFX_SORPTION_LANGMUIR
    d = _deck("   A(aq)\n    FREUNDLICH_N 2.\n    DISTRIBUTION_COEFFICIENT 2.\n    TYPE FREUNDLICH\n    /\n")
    assert d.chemistry.isotherms[0].type == abi.PFX_SORPTION_FREUNDLICH
    with pytest.raises(DeckError):
        _deck("   A(aq)\n    TYPE linear\n    /\n")     # the type word is compared as written


def test_deck_units_carry_over_to_later_isotherms():
    # the units word is kept across reactions (reaction.F90:142, 180, 563)
    d = _deck("   A(aq)\n    DISTRIBUTION_COEFFICIENT 2. g/m^3\n    /\n"
              "   B(aq)\n    DISTRIBUTION_COEFFICIENT 5.\n    /\n")
    a, b = d.chemistry.isotherms
    assert abs(a.kd - 2e-3) <= 1e-18 and abs(b.kd - 5e-3) <= 1e-18


# ---------------------------------------------------------------- oracle


def test_mineral_scaled_kd_is_the_decks_250_kg_per_m3():
    S_min, _ = sorbed(tracer(abi.PFX_SORPTION_LINEAR, 0.125786163522, mineral=True), M)
    S_lin, _ = sorbed(tracer(abi.PFX_SORPTION_LINEAR, 250.0), M)
    assert np.all(np.abs(S_min / S_lin - 1.0) <= 1e-11)
    # retardation 2: as much sorbed as dissolved (phi s rho_w m)
    assert np.all(np.abs(S_lin / (0.25 * 1000.0 * M) - 1.0) <= 1e-14)
    # the volume fraction is only a factor
    S_half, _ = sorbed(tracer(abi.PFX_SORPTION_LINEAR, 0.125786163522, mineral=True), M, vf=0.5)
    assert np.allclose(S_half, 0.5 * S_min, rtol=1e-14, atol=0.0)


def test_langmuir_limits():
    kd, b = 1e4, 0.3
    S, _ = sorbed(tracer(abi.PFX_SORPTION_LANGMUIR, kd, b=b), M)
    x = kd * M
    hi, lo = x > 1e4, x < 1e-4
    assert np.all(np.abs(S[hi] / b - 1.0) <= 2.0 / x[hi])          # saturation
    assert np.all(np.abs(S[lo] / (b * x[lo]) - 1.0) <= 2.0 * x[lo])   # low coverage
    assert np.all(np.diff(S) > 0.0)


def test_freundlich_homogeneity_and_linear_case():
    nf = 2.3
    t = tracer(abi.PFX_SORPTION_FREUNDLICH, 7.0, n=nf)
    S1, _ = sorbed(t, M)
    S2, _ = sorbed(t, 10.0 * M)
    assert np.all(np.abs(S2 / S1 - 10.0 ** (1.0 / nf)) <= 1e-13)
    Sf, _ = sorbed(tracer(abi.PFX_SORPTION_FREUNDLICH, 7.0, n=1.0), M)
    Sl, _ = sorbed(tracer(abi.PFX_SORPTION_LINEAR, 7.0), M)
    assert np.allclose(Sf, Sl, rtol=1e-15, atol=0.0)


@pytest.mark.parametrize("case", [(abi.PFX_SORPTION_LINEAR, 40.0, 0.0, 0.0, False),
                                  (abi.PFX_SORPTION_LANGMUIR, 3e3, 0.2, 0.0, False),
                                  (abi.PFX_SORPTION_FREUNDLICH, 5.0, 0.0, 1.7, False),
                                  (abi.PFX_SORPTION_LINEAR, 0.4, 0.0, 0.0, True)],
                         ids=["linear", "langmuir", "freundlich", "mineral"])
def test_derivative_is_the_derivative(case):
    kind, kd, b, n, mineral = case
    t = tracer(kind, kd, b=b, n=n, mineral=mineral)
    m = np.logspace(-8, -2, 7)
    _, dS = sorbed(t, m)
    h = 1e-5 * m
    Sp, _ = sorbed(t, m + h)
    Sm, _ = sorbed(t, m - h)
    fd = (Sp - Sm) / (2.0 * h)
    assert np.all(np.abs(dS - fd) <= 1e-7 * np.abs(fd))


def _bisect(f, lo, hi, iters=200):
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0.0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("case", [(abi.PFX_SORPTION_LINEAR, 40.0, 0.0, 0.0, False),
                                  (abi.PFX_SORPTION_LANGMUIR, 3e3, 0.2, 0.0, False),
                                  (abi.PFX_SORPTION_FREUNDLICH, 5.0, 0.0, 1.7, False),
                                  (abi.PFX_SORPTION_LINEAR, 0.4, 0.0, 0.0, True)],
                         ids=["linear", "langmuir", "freundlich", "mineral"])
def test_reaction_step_conserves_the_sorbing_tracer(case):
    # a cell whose old state has the tracer all dissolved (no sorbed amount):
    # the step's free ion solves phi s rho_w V m + S(m) V = phi s 1000 V T_old
    kind, kd, b, n, mineral = case
    t = tracer(kind, kd, b=b, n=n, mineral=mineral)
    phi, den = 0.25, 997.16
    tot = np.array([1e-7, 1e-5, 1e-3])
    c = orc.HostCells.allocate(t, len(tot))
    c["total"][:, 0] = tot
    c["total_sorb"][:] = 0.0
    c["pri_molal"][:, 0] = tot * 1000.0 / den
    if mineral:
        c["mnrl_volfrac"][:, 0] = 0.3
    o = abi.default_opts(dt=3600.0, den_kg=den, porosity=phi)
    rc, _ = orc.react_batch(t, c, o)
    assert rc == 0
    for i, T in enumerate(tot):
        def excess(m):
            s, _ = sorbed(t, np.array([m]), porosity=phi, den=den, vf=0.3)
            return phi * den * m + s[0] - phi * 1000.0 * T
        want = _bisect(excess, 0.0, T * 1000.0 / den)
        assert abs(c["pri_molal"][i, 0] / want - 1.0) <= 1e-9, (i, c["pri_molal"][i, 0], want)


@pytest.mark.parametrize("kind", [abi.PFX_SORPTION_LINEAR, abi.PFX_SORPTION_LANGMUIR, abi.PFX_SORPTION_FREUNDLICH])
def test_total_sorbed_constraint_inverts_the_isotherm(kind):
    kd, b, n = {abi.PFX_SORPTION_LINEAR: (40.0, 0.0, 0.0), abi.PFX_SORPTION_LANGMUIR: (3e3, 0.2, 0.0),
                abi.PFX_SORPTION_FREUNDLICH: (5.0, 0.0, 1.7)}[kind]
    t = tracer(kind, kd, b=b, n=n)
    target = np.array([1e-6, 1e-4, 1e-2 if kind != abi.PFX_SORPTION_LANGMUIR else 0.1])
    c = orc.HostCells.allocate(t, len(target))
    rc, its = orc.equilibrate(t, [abi.PFX_CONSTRAINT_TOTAL_SORB], target.reshape(-1, 1), c,
                              synth.react_opts(porosity=0.25))
    assert rc == 0 and np.all(its > 0)
    m = c["pri_molal"][:, 0]
    # the inverse isotherms: S / K_D, S / (K_D (b - S)), (S / K_D)^n
    want = {abi.PFX_SORPTION_LINEAR: target / kd, abi.PFX_SORPTION_LANGMUIR: target / (kd * (b - target)),
            abi.PFX_SORPTION_FREUNDLICH: (target / kd) ** n}[kind]
    assert np.all(np.abs(m / want - 1.0) <= 1e-10)
    assert np.all(np.abs(c["total_sorb"][:, 0] / target - 1.0) <= 1e-12)


def test_synthetic_isotherm_network_has_every_type():
    t = synth.with_isotherms(synth.random_network(15, 50, 3, ionx=True))
    assert sorted(t.kd_type.tolist()) == [1, 1, 2, 3] and t.kd_mineral.tolist().count(0) == 1
    assert len(set(t.kd_specid.tolist())) == 4 and 0 not in t.kd_specid
