"""Seeded synthetic chemistry and cell inputs for the batched reaction step.

Workloads (sizes as named in the benchmark description):
  * ``carbonate_column``: 5 primaries (H+, HCO3-, Ca++, Mg++, Cl-), 8 complexes
    and kinetic calcite from the hanford database, 25 C, Debye-Hueckel.
  * ``random_network``: N primaries (H+ first), M complexes with stoichiometry
    in {-2..2}\\{0} on 2-4 primaries (formation-biased: P(nu) = .05/.15/.5/.3
    for -2/-1/1/2) and log K ~ N(0, 6) at 25 C, accepted when the complex
    stays below 10 mol/kg over free ions in [1e-9, 1e-2] molal, with a seeded
    van 't Hoff temperature dependence fitted to the 5-term form, K kinetic
    minerals (3-5 primaries, k log-uniform [1e-12, 1e-8] mol/m^2/s) and one
    ion-exchange site (CEC 750 eq/m^3) over the cationic primaries.

Every random draw comes from a counter-based generator (numpy Philox) keyed
by (workload, field), so CPU and GPU see bit-identical inputs.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import abi
from .database import (ReactionTables, debye_params, logk_eval, logk_fit, water_density_ifc67)

SEED = 20261202
DB_TEMPS = [0.0, 25.0, 60.0, 100.0, 150.0, 200.0, 250.0, 300.0]

FIELDS = {"table": 1, "totals": 2, "ph": 3, "temp": 4, "vf": 5, "guess": 6, "perturb": 7, "perm": 8, "isotherms": 9}


def rng(workload_id: int, field: str, seed: int = SEED) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[seed, (workload_id << 16) | FIELDS[field]]))


def random_network(N: int = 15, M: int = 50, K: int = 3, ionx: bool = True, seed: int = SEED,
                   workload_id: int = 2, act_freq: int = abi.PFX_ACT_TIMESTEP) -> ReactionTables:
    g = rng(workload_id, "table", seed)
    # primaries: H+ first, then charges in {-2,-1,1,2}
    Z = np.empty(N)
    Z[0] = 1.0
    Z[1:] = g.choice([-2.0, -1.0, 1.0, 2.0], size=N - 1, p=[0.2, 0.3, 0.3, 0.2])
    if not np.any(Z[1:] > 0):
        Z[1] = 2.0
    a0 = g.choice([3.0, 4.0, 4.5, 5.0, 6.0, 8.0, 9.0], size=N)
    a0[0] = 9.0
    S = 4
    spec = np.zeros((M, S + 1), dtype=np.int32)
    nu = np.zeros((M, S))
    logK25 = np.zeros(M)
    coefs = np.zeros((M, 5))
    cZ = np.zeros(M)
    lo_ln_m, hi_ln_m = np.log(1e-9), np.log(1e-2)
    for i in range(M):
        while True:
            ns = int(g.integers(2, 5))
            sp = np.sort(g.choice(N, size=ns, replace=False))
            st = g.choice([-2.0, -1.0, 1.0, 2.0], size=ns, p=[0.05, 0.15, 0.5, 0.3])
            lk = float(g.normal(0.0, 6.0))
            # acceptance: the complex stays below 10 mol/kg for every free-ion
            # state in [1e-9, 1e-2] molal, i.e. from the default guess to the
            # most concentrated water of the workload
            ln_c = -lk * np.log(10.0) + float(np.sum(np.maximum(st * lo_ln_m, st * hi_ln_m)))
            if ln_c < np.log(10.0):
                break
        spec[i, 0] = ns
        spec[i, 1:1 + ns] = sp
        nu[i, :ns] = st
        logK25[i] = lk
        cZ[i] = float(np.sum(st * Z[sp]))
        dH = float(g.normal(0.0, 30e3))  # J/mol, van 't Hoff slope
        lks = [lk + dH / (8.31446 * np.log(10.0)) * (1.0 / 298.15 - 1.0 / (t + 273.15)) for t in DB_TEMPS]
        coefs[i] = logk_fit(DB_TEMPS, lks)
    ca0 = np.full(M, 4.0)
    # kinetic minerals
    mspec = np.zeros((K, S + 1), dtype=np.int32)
    mnu = np.zeros((K, S))
    mlogK = np.zeros(K)
    mcoef = np.zeros((K, 5))
    mk = np.zeros(K)
    for k in range(K):
        ns = int(g.integers(3, 5))
        sp = np.sort(g.choice(N, size=ns, replace=False))
        st = g.choice([-2.0, -1.0, 1.0, 2.0], size=ns)
        mspec[k, 0] = ns
        mspec[k, 1:1 + ns] = sp
        mnu[k, :ns] = st
        # saturation near typical states: log K = ion activity product at 1e-4
        mlogK[k] = float(np.sum(st)) * -4.0 + float(g.normal(0.0, 1.0))
        lks = [mlogK[k] + float(g.normal(0, 0.3)) * (t - 25.0) / 100.0 for t in DB_TEMPS]
        mcoef[k] = logk_fit(DB_TEMPS, lks)
        mk[k] = 10.0 ** g.uniform(-12.0, -8.0)
    # ion exchange over the cationic primaries, reference = first cation
    if ionx:
        cats = [j for j in range(N) if Z[j] > 0][:8]
        R = 1
        ionx_id = np.zeros((1, len(cats) + 1), dtype=np.int32)
        ionx_id[0, 0] = len(cats)
        ionx_id[0, 1:] = cats
        ionx_k = np.zeros((1, len(cats)))
        ionx_k[0, 0] = 1.0
        ionx_k[0, 1:] = 10.0 ** g.uniform(-1.0, 1.0, size=len(cats) - 1)
        cec = np.array([750.0])
        to_surf = np.array([-1], dtype=np.int32)
    else:
        R = 0
        ionx_id = np.zeros((0, 2), dtype=np.int32)
        ionx_k = np.zeros((0, 1))
        cec = np.zeros(0)
        to_surf = np.zeros(0, dtype=np.int32)
    names = ["H+"] + [f"P{j}" for j in range(1, N)]
    return ReactionTables(
        primary_names=names, complex_names=[f"X{i}" for i in range(M)], mineral_names=[f"Min{k}" for k in range(K)],
        Z=Z, a0=a0, cplx_specid=spec, cplx_stoich=nu, cplx_h2ostoich=np.zeros(M), cplx_Z=cZ, cplx_a0=ca0,
        cplx_logK=logK25, cplx_logKcoef=coefs, mnrl_specid=mspec, mnrl_stoich=mnu, mnrl_h2ostoich=np.zeros(K),
        mnrl_logK=mlogK, mnrl_logKcoef=mcoef, mnrl_rate_constant=mk, mnrl_molar_vol=np.full(K, 3.0e-5),
        mnrl_activation_energy=np.zeros(K), mnrl_affinity_threshold=np.zeros(K), mnrl_rate_limiter=np.zeros(K),
        ionx_cationid=ionx_id, ionx_k=ionx_k, ionx_CEC=cec, ionx_to_surf=to_surf,
        debye=debye_params(25.0), use_log_formulation=1, act_freq=act_freq, act_alg=abi.PFX_ACT_ALG_LAG)


def with_isotherms(t: ReactionTables, seed: int = SEED, workload_id: int = 2,
                   soil_particle_density: float = 2650.0) -> ReactionTables:
    """The network with isotherm (K_D) sorption on four non-H+ primaries:
    linear, Langmuir, Freundlich, and linear scaled by kinetic mineral 0's
    volume fraction (K_D in mL/g).  Sized so that the sorbed and aqueous
    amounts of a species are comparable over 1e-9..1e-2 molal."""
    from dataclasses import replace
    g = rng(workload_id, "isotherms", seed)
    sp = np.sort(g.choice(np.arange(1, t.N), size=4, replace=False)).astype(np.int32)
    kd = np.array([10.0 ** g.uniform(1.0, 3.0), 10.0 ** g.uniform(2.0, 5.0), 10.0 ** g.uniform(0.0, 2.0),
                   10.0 ** g.uniform(-2.0, 0.0)])
    return replace(
        t, kd_specid=sp,
        kd_type=np.array([abi.PFX_SORPTION_LINEAR, abi.PFX_SORPTION_LANGMUIR, abi.PFX_SORPTION_FREUNDLICH,
                          abi.PFX_SORPTION_LINEAR], dtype=np.int32),
        kd_coef=kd, kd_langmuir_b=np.array([0.0, 10.0 ** g.uniform(-3.0, 0.0), 0.0, 0.0]),
        kd_freundlich_n=np.array([0.0, 0.0, g.uniform(1.2, 3.0), 0.0]),
        kd_mineral=np.array([-1, -1, -1, 0 if t.K > 0 else -1], dtype=np.int32),
        soil_particle_density=soil_particle_density, _desc=None)


@dataclass
class CellInputs:
    """Constraint definition of each cell: totals (mol/kg) of every primary
    but H+, pH, temperature and mineral volume fractions."""
    totals: np.ndarray      # [n, N]; column 0 (H+) unused
    ph: np.ndarray          # [n]
    temp: np.ndarray        # [n] deg C
    vf: np.ndarray          # [n, K]
    area: np.ndarray        # [n, K]


def random_cells(t: ReactionTables, n: int, seed: int = SEED, workload_id: int = 2,
                 total_range=(1e-6, 1e-2), ph_range=(5.0, 9.0), temp_range=(10.0, 60.0),
                 vf_range=(0.01, 0.1)) -> CellInputs:
    lo, hi = np.log10(total_range[0]), np.log10(total_range[1])
    totals = 10.0 ** rng(workload_id, "totals", seed).uniform(lo, hi, size=(n, t.N))
    ph = rng(workload_id, "ph", seed).uniform(*ph_range, size=n)
    temp = rng(workload_id, "temp", seed).uniform(*temp_range, size=n)
    vf = rng(workload_id, "vf", seed).uniform(*vf_range, size=(n, max(t.K, 1)))
    area = np.ones((n, max(t.K, 1)))
    return CellInputs(totals, ph, temp, vf, area)


def host_state_arrays(t: ReactionTables, n: int, with_temperature: bool) -> Dict[str, Optional[np.ndarray]]:
    """Empty pfx_cells arrays (cell-major) for n cells."""
    N, M, K, R = t.N, t.M, t.K, t.R
    z = lambda *s: np.zeros(s, dtype=np.float64)  # noqa: E731
    return {
        "total": z(n, N), "total_sorb": z(n, N) if t.has_sorption else None, "pri_molal": np.full((n, N), 1e-9),
        "pri_act_coef": np.ones((n, N)), "sec_act_coef": np.ones((n, max(M, 1))), "sec_molal": z(n, max(M, 1)),
        "ionx_ref_sorbed": np.full((n, max(R, 1)), 1e-9), "mnrl_volfrac": z(n, max(K, 1)),
        "mnrl_area": z(n, max(K, 1)), "mnrl_rate": z(n, max(K, 1)),
        "temperature": (z(n) + 25.0) if with_temperature else None,
        "porosity": None, "saturation": None, "volume": None, "den_kg": None,
        "guess": z(n, N), "num_iterations": np.zeros(n, dtype=np.int32),
        "status": np.zeros(n, dtype=np.int32), "rel_res": z(n),
    }


DEN_25C = water_density_ifc67(25.0, 101325.0)
STEP_DT = 8640.0   # 0.1 d


# step test of the reaction step's convergence: besides rel_res <= 1e-10,
# the Newton step at the accepted iterate, enlarged by the observed
# contraction, stays below 1e-11 relative (free ions to ~1e-11)
REL_CHANGE_TOL = 1e-11


def react_opts(**kw) -> abi.ReactOpts:
    base = dict(dt=STEP_DT, rel_tol=1e-10, max_iterations=150, extra_iterations=0,
                guess_mode=abi.PFX_GUESS_DEFAULT, default_guess=1e-9, porosity=0.25, saturation=1.0,
                volume=1.0, den_kg=DEN_25C, rel_change_tol=REL_CHANGE_TOL)
    base.update(kw)
    return abi.default_opts(**base)


def logk_at(t: ReactionTables, temp_c: np.ndarray) -> np.ndarray:
    """per-cell log K of every complex (host check of the in-kernel fit)."""
    return logk_eval(t.cplx_logKcoef[None, :, :], np.asarray(temp_c)[:, None])
