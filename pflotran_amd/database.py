"""Reaction tables: the flattened ``reaction_type`` handed to the C-ABI.

Host-side setup that a PFLOTRAN Fortran host would do in its own
``DatabaseRead``/``BasisInit`` (src/pflotran/reaction_database.F90:25, 772).
This module reads PFLOTRAN's ASCII thermodynamic database format, writes
every reaction a deck uses in the deck's primary species (Basis), and builds
the numpy arrays of ``pfx_reaction_desc`` (include/pfx_chem.h).
"""
from __future__ import annotations

import ctypes as C
import math
import shlex
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import abi

# pflotran_constants.F90:48 (truncated on purpose; the kernels use the same)
LOG_TO_LN = 2.30258509299
H2O_NAMES = ("H2O",)


# --------------------------------------------------------------------------
# database file


@dataclass
class DbSpecies:
    name: str
    a0: float
    Z: float
    mw: float


@dataclass
class DbReaction:
    name: str
    species: List[Tuple[float, str]]
    logK: List[float]
    a0: float = 0.0
    Z: float = 0.0
    mw: float = 0.0
    molar_volume: float = 0.0  # minerals: cm^3/mol as written


@dataclass
class Database:
    temperatures: List[float]
    primary: Dict[str, DbSpecies] = field(default_factory=dict)
    complexes: Dict[str, DbReaction] = field(default_factory=dict)
    gases: Dict[str, DbReaction] = field(default_factory=dict)
    minerals: Dict[str, DbReaction] = field(default_factory=dict)


def _num(tok: str) -> float:
    return float(tok.lower().replace("d", "e"))


def read_database(path: str) -> Database:
    """Parse the block layout of reaction_database.F90:113-420: temperatures,
    then primary species | complexes | gases | minerals | surface complexes,
    each section closed by a 'null' entry."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    head = shlex.split(lines[0])
    ntemp = int(head[1])
    db = Database(temperatures=[_num(t) for t in head[2:2 + ntemp]])
    section = 0
    for ln in lines[1:]:
        toks = shlex.split(ln)
        if not toks:
            continue
        name = toks[0]
        if name == "null":
            section += 1
            if section >= 5:
                break
            continue
        try:
            if section == 0:
                db.primary[name] = DbSpecies(name, _num(toks[1]), _num(toks[2]), _num(toks[3]))
            elif section == 1:
                n = int(toks[1])
                sp = [(_num(toks[2 + 2 * i]), toks[3 + 2 * i]) for i in range(n)]
                p = 2 + 2 * n
                lk = [_num(t) for t in toks[p:p + ntemp]]
                a0, Z, mw = (_num(t) for t in toks[p + ntemp:p + ntemp + 3])
                db.complexes[name] = DbReaction(name, sp, lk, a0, Z, mw)
            elif section in (2, 3):
                mv = _num(toks[1])
                n = int(toks[2])
                sp = [(_num(toks[3 + 2 * i]), toks[4 + 2 * i]) for i in range(n)]
                p = 3 + 2 * n
                lk = [_num(t) for t in toks[p:p + ntemp]]
                mw = _num(toks[p + ntemp])
                tgt = db.gases if section == 2 else db.minerals
                tgt[name] = DbReaction(name, sp, lk, mw=mw, molar_volume=mv)
        except (IndexError, ValueError):
            # entries this build never selects (e.g. malformed test species)
            continue
    return db


# --------------------------------------------------------------------------
# basis: every reaction written in the deck's primary species


def _flatten(db: Database, name: str, memo: Dict, stack=(), keep=None):
    """(composition over the database's primary species, H2O coefficient,
    log K per database temperature, NaN where missing) of a species: a
    database primary is itself; a complex, gas or mineral is its reaction
    with every non-primary species in it replaced by its own reaction
    (recursively), so log a_X = sum_k c_k log a_k + h log a_w - log K.  A
    database primary outside `keep` (the chosen primaries) that is also
    defined by a gas or complex reaction of the same name (O2(g) in
    hanford.dat, a primary and a gas O2(g) = O2(aq)) is replaced by that
    reaction."""
    if name in memo:
        return memo[name]
    nt = len(db.temperatures)
    has_rxn = name in db.gases or name in db.complexes
    if name in db.primary and (keep is None or name in keep or not has_rxn or name in stack):
        out = ({name: 1.0}, 0.0, np.zeros(nt))
    else:
        r = db.gases.get(name) or db.complexes.get(name) or db.minerals.get(name)
        if r is None:
            raise ValueError(f"species {name} is not in the database")
        if name in stack:
            raise ValueError(f"species {name}: circular reaction")
        comp: Dict[str, float] = {}
        h2o = 0.0
        lk = np.array([np.nan if abs(v - 500.0) < 1e-10 else v for v in r.logK], dtype=np.float64)
        for st, sp in r.species:
            if sp in H2O_NAMES:
                h2o += st
                continue
            c, h, l = _flatten(db, sp, memo, stack + (name,), keep)
            for k, v in c.items():
                comp[k] = comp.get(k, 0.0) + st * v
            h2o += st * h
            lk = lk + st * l
        out = (comp, h2o, lk)
    memo[name] = out
    return out


class Basis:
    """The deck's primary species as the basis of every reaction, the role
    of BasisInit's basis handling (reaction_database.F90:1246-1380).  A
    reaction whose species are all primaries (or H2O) is taken as written.
    Otherwise each species is reduced to the database's own primaries
    (_flatten), then written in the chosen primaries by solving
    sum_p x_p comp(p) = comp(X); its log K becomes logK(X) - sum_p x_p
    logK(p) (0 for a database primary), per temperature, missing (500) where
    a term is.  A primary that is a database complex takes its charge, ion
    size and molar mass from there."""

    def __init__(self, db: Database, primaries: Sequence[str]):
        self.db = db
        self.prim = list(primaries)
        self.pset = set(self.prim)
        self._memo: Dict = {}
        self._pc = None
        for p in self.prim:
            if p not in db.primary and p not in db.complexes:
                raise ValueError(f"primary species {p} is not in the database")

    def primary(self, p: str) -> DbSpecies:
        if p in self.db.primary:
            return self.db.primary[p]
        r = self.db.complexes[p]
        return DbSpecies(p, r.a0, r.Z, r.mw)

    def reaction(self, name: str, section: Dict[str, DbReaction]) -> DbReaction:
        r = section.get(name)
        if r is None and name in self.db.primary:
            # a database primary made secondary by the deck's choice of basis
            sp = self.db.primary[name]
            r = DbReaction(name, [(1.0, name)], [0.0] * len(self.db.temperatures), sp.a0, sp.Z, sp.mw)
        elif r is None:
            raise ValueError(f"species {name} is not in the database")
        if all(sp in self.pset or sp in H2O_NAMES for _, sp in r.species):
            return r
        if self._pc is None:
            self._pc = [_flatten(self.db, p, self._memo, keep=self.pset) for p in self.prim]
            self._basis = sorted({k for c, _, _ in self._pc for k in c})
            self._A = np.array([[c.get(k, 0.0) for k in self._basis] for c, _, _ in self._pc], dtype=np.float64)
        c, h, lk = _flatten(self.db, name, self._memo, keep=self.pset)
        missing = [k for k, v in c.items() if abs(v) > 1e-12 and k not in self._basis]
        if missing:
            raise ValueError(f"{name}: species {missing} cannot be written in the primary species")
        cv = np.array([c.get(k, 0.0) for k in self._basis], dtype=np.float64)
        x, *_ = np.linalg.lstsq(self._A.T, cv, rcond=None)
        if np.max(np.abs(self._A.T @ x - cv), initial=0.0) > 1e-9:
            raise ValueError(f"{name}: reaction cannot be written in the primary species")
        x = np.where(np.abs(x - np.round(x)) < 1e-12, np.round(x), x)
        lk2, h2 = lk.copy(), h
        for xp, (_, hp, lkp) in zip(x, self._pc):
            if xp != 0.0:
                lk2 = lk2 - xp * lkp
                h2 = h2 - xp * hp
        species = [(float(xp), p) for xp, p in zip(x, self.prim) if xp != 0.0]
        if abs(h2) > 1e-12:
            species.append((float(h2), "H2O"))
        import dataclasses
        return dataclasses.replace(r, species=species,
                                   logK=[500.0 if not np.isfinite(v) else float(v) for v in lk2])


# --------------------------------------------------------------------------
# closed-form setup helpers


def debye_params(tref: float = 25.0, use_bdot: bool = True) -> Tuple[float, float, float]:
    """Debye-Hueckel A, B, Bdot at the reference temperature
    (reaction_database.F90:896-990): piecewise-linear in fixed tables."""
    T = [0, 25, 60, 100, 150, 200, 250, 300, 350]
    a = [0.4939, 0.5114, 0.5465, 0.5995, 0.6855, 0.7994, 0.9593, 1.2180, 1.2180]
    b = [0.3253, 0.3288, 0.3346, 0.3421, 0.3525, 0.3639, 0.3766, 0.3925, 0.3925]
    d = [0.0374, 0.0410, 0.0440, 0.0460, 0.0470, 0.0470, 0.0340, 0.0000, 0.0000]
    if tref <= 0.01:
        A, B, Bd = a[0], b[0], d[0]
    elif tref > 350.0:
        A, B, Bd = a[8], b[8], d[8]
    else:
        k = 1
        while k < 8 and tref > T[k]:
            k += 1
        w = (tref - T[k - 1]) / (T[k] - T[k - 1])
        A = a[k - 1] + w * (a[k] - a[k - 1])
        B = b[k - 1] + w * (b[k] - b[k - 1])
        Bd = d[k - 1] + w * (d[k] - d[k - 1])
    return A, B, (Bd if use_bdot else 0.0)


def _crout_solve(a: List[List[float]], b: List[float]) -> List[float]:
    """Crout LU with implicit partial pivoting and substitution, in the
    reference's operation order (utility.F90:393-523), for the small normal
    equations of the log K fit; the fit is ill-conditioned enough that a
    different elimination order moves log K by ~1e-8."""
    n = len(b)
    A = [row[:] for row in a]
    vv = [1.0 / max(abs(x) for x in row) for row in A]
    indx = [0] * n
    for j in range(n):
        for i in range(j):
            s = A[i][j]
            for k in range(i):
                s -= A[i][k] * A[k][j]
            A[i][j] = s
        aamax, imax = 0.0, j
        for i in range(j, n):
            s = A[i][j]
            for k in range(j):
                s -= A[i][k] * A[k][j]
            A[i][j] = s
            dum = vv[i] * abs(s)
            if dum >= aamax:
                imax, aamax = i, dum
        if j != imax:
            A[imax], A[j] = A[j], A[imax]
            vv[imax] = vv[j]
        indx[j] = imax
        if A[j][j] == 0.0:
            A[j][j] = 1e-20
        if j != n - 1:
            dum = 1.0 / A[j][j]
            for i in range(j + 1, n):
                A[i][j] *= dum
    x = list(b)
    ii = -1
    for i in range(n):
        ll = indx[i]
        s = x[ll]
        x[ll] = x[i]
        if ii >= 0:
            for j in range(ii, i):
                s -= A[i][j] * x[j]
        elif s != 0.0:
            ii = i
        x[i] = s
    for i in range(n - 1, -1, -1):
        s = x[i]
        for j in range(i + 1, n):
            s -= A[i][j] * x[j]
        x[i] = s / A[i][i]
    return x


def logk_fit(temps: Sequence[float], logK: Sequence[float]) -> np.ndarray:
    """5-term Maier-Kelly least-squares fit of log K(T) over the database
    temperatures, entries of 500 are missing (reaction_aux.F90:1336-1408):
    normal equations accumulated and solved in the reference's order."""
    vec = []
    for t in temps:
        tk = float(t) + 273.15
        vec.append([math.log(tk), 1.0, tk, 1.0 / tk, 1.0 / (tk * tk)])
    use = [abs(float(k) - 500.0) >= 1e-10 for k in logK]
    if sum(use) < 5:
        # under-determined (e.g. one tabulated temperature): minimum-norm fit;
        # only used for non-isothermal runs, which need >= 5 temperatures
        V = np.array([v for v, u in zip(vec, use) if u])
        return np.linalg.lstsq(V, np.asarray(logK, dtype=np.float64)[np.array(use)], rcond=None)[0]
    coefs = []
    for j in range(5):
        s = 0.0
        for i in range(len(temps)):
            if use[i]:
                s += vec[i][j] * float(logK[i])
        coefs.append(s)
    a = [[0.0] * 5 for _ in range(5)]
    for j in range(5):
        for k in range(j, 5):
            s = 0.0
            for i in range(len(temps)):
                if use[i]:
                    s += vec[i][j] * vec[i][k]
            a[j][k] = s
            a[k][j] = s
    return np.array(_crout_solve(a, coefs))


def logk_eval(coefs: np.ndarray, temp_c) -> np.ndarray:
    """reaction_aux.F90:1461-1488 ReactionInterpolateLogK"""
    tk = np.asarray(temp_c, dtype=np.float64) + 273.15
    c = np.asarray(coefs)
    return c[..., 0] * np.log(tk) + c[..., 1] + c[..., 2] * tk + c[..., 3] / tk + c[..., 4] / (tk * tk)


def water_density_ifc67(t: float = 25.0, p: float = 101325.0) -> float:
    """IFC-67 liquid water density in kg/m^3 (eos_water.F90:876-940), the
    reference water density of a transport-only run (factory_subsurface.F90:995)."""
    aa = [6.824687741e03, -5.422063673e02, -2.096666205e04, 3.941286787e04,
          -6.733277739e04, 9.902381028e04, -1.093911774e05, 8.590841667e04,
          -4.511168742e04, 1.418138926e04, -2.017271113e03, 7.982692717e00,
          -2.616571843e-2, 1.522411790e-3, 2.284279054e-2, 2.421647003e02,
          1.269716088e-10, 2.074838328e-7, 2.174020350e-8, 1.105710498e-9,
          1.293441934e01, 1.308119072e-5, 6.047626338e-14]
    a1, a2, a3, a4 = 8.438375405e-1, 5.362162162e-4, 1.72, 7.342278489e-2
    a5, a6, a7, a8 = 4.975858870e-2, 6.537154300e-1, 1.15e-6, 1.5108e-5
    a9, a10, a11, a12 = 1.4188e-1, 7.002753165e00, 2.995284926e-4, 2.04e-1
    theta = (t + 273.15) * (1.0 / 647.3)
    th2 = theta * theta
    th18 = theta ** 18.0
    th20 = th18 * th2
    beta = p * (1.0 / 22.064e6)
    b2 = beta * beta
    yy = 1.0 - a1 * th2 - a2 * theta ** (-6.0)
    xx = a3 * yy * yy - 2.0 * (a4 * theta - a5 * beta)
    xx = math.sqrt(xx) if xx > 0 else 1e-6
    zz = yy + xx
    u1 = aa[11] * a5 * zz ** (-5.0 / 17.0)
    u2 = 1.0 / (a8 + theta ** 11.0)
    u3 = aa[17] + (2.0 * aa[18] + 3.0 * aa[19] * beta) * beta
    u4 = 1.0 / (a7 + th18 * theta)
    u5 = (a10 + beta) ** (-4.0)
    u6 = a11 - 3.0 * u5
    u7 = aa[20] * th18 * (a9 + th2)
    u8 = aa[15] * (a6 - theta) ** 9.0
    vr = (u1 + aa[12] + theta * (aa[13] + aa[14] * theta) + u8 * (a6 - theta) + aa[16] * u4
          - u2 * u3 - u6 * u7 + (3.0 * aa[21] * (a12 - theta) + 4.0 * aa[22] * beta / th20) * b2)
    return 1.0 / (vr * 0.00317)


# --------------------------------------------------------------------------
# flattened tables


@dataclass
class PrefactorSpecies:
    """One factor a^alpha / (1 + atten a^beta) of a rate prefactor
    (ts_prefactor_species_type, reaction_mineral_aux.F90; defaults 0 as
    TSPrefactorSpeciesCreate sets them)."""
    name: str
    alpha: float = 0.0
    beta: float = 0.0
    atten: float = 0.0


@dataclass
class Prefactor:
    """A parallel rate term of a kinetic mineral (PREFACTOR block): its rate
    constant and activation energy default to the mineral's own
    (reaction_mineral.F90:343-360)."""
    species: List[PrefactorSpecies]
    rate: Optional[float] = None               # mol/m^2/s
    activation_energy: Optional[float] = None  # J/mol


@dataclass
class MineralKinetics:
    rate_constant: float           # mol/m^2/s
    activation_energy: float = 0.0  # J/mol
    affinity_threshold: float = 0.0
    rate_limiter: float = 0.0
    affinity_power: Optional[float] = None   # beta (AFFINITY_POWER)
    temkin: Optional[float] = None           # sigma (TEMKIN_CONSTANT)
    prefactors: List[Prefactor] = field(default_factory=list)
    min_scale_factor: Optional[float] = None  # f (MIN_SCALE_FACTOR)


def _optional_per_mineral(vals):
    """An optional per-mineral coefficient: absent (None) unless a mineral
    sets it, then set for all with 1 for those that do not."""
    if all(v is None for v in vals):
        return None
    return np.array([1.0 if v is None else float(v) for v in vals], dtype=np.float64)


@dataclass
class IonExchangeRxn:
    cations: List[Tuple[str, float]]   # reference cation first, k=1
    CEC: float                         # eq/m^3 bulk
    mineral: Optional[str] = None


@dataclass
class KDRxn:
    """An isotherm reaction of SORPTION / ISOTHERM_REACTIONS (kd_rxn_type,
    reaction_aux.F90:80-90).  K_D is in kg water/m^3 bulk, or with a
    mineral in mL water/g soil (L/kg) scaled by the soil density and the
    mineral's volume fraction."""
    species: str
    kd: float
    type: int = abi.PFX_SORPTION_LINEAR
    langmuir_b: float = 0.0
    freundlich_n: float = 0.0
    mineral: Optional[str] = None


@dataclass
class ReactionTables:
    """numpy form of pfx_reaction_desc; `desc()` returns the ctypes struct
    (the arrays must outlive it: they are kept on this object)."""
    primary_names: List[str]
    complex_names: List[str]
    mineral_names: List[str]
    Z: np.ndarray
    a0: np.ndarray
    cplx_specid: np.ndarray      # [M][S+1] int32
    cplx_stoich: np.ndarray      # [M][S]
    cplx_h2ostoich: np.ndarray
    cplx_Z: np.ndarray
    cplx_a0: np.ndarray
    cplx_logK: np.ndarray
    cplx_logKcoef: Optional[np.ndarray]
    mnrl_specid: np.ndarray
    mnrl_stoich: np.ndarray
    mnrl_h2ostoich: np.ndarray
    mnrl_logK: np.ndarray
    mnrl_logKcoef: Optional[np.ndarray]
    mnrl_rate_constant: np.ndarray
    mnrl_molar_vol: np.ndarray
    mnrl_activation_energy: np.ndarray
    mnrl_affinity_threshold: np.ndarray
    mnrl_rate_limiter: np.ndarray
    ionx_cationid: np.ndarray    # [R][C+1]
    ionx_k: np.ndarray           # [R][C]
    ionx_CEC: np.ndarray
    ionx_to_surf: np.ndarray
    debye: Tuple[float, float, float] = (0.5114, 0.3288, 0.0410)
    use_log_formulation: int = 1
    act_freq: int = abi.PFX_ACT_OFF
    act_alg: int = abi.PFX_ACT_ALG_LAG
    numerical_jacobian: int = 0
    max_dlnC: float = 5.0
    max_rel_change_tol: float = 1e-6
    max_residual_tol: float = 1e-12
    # AFFINITY_POWER / TEMKIN_CONSTANT: allocated for every kinetic mineral
    # when any mineral sets one, 1 for the others (reaction_database.F90
    # tables built in BasisInit); None when no mineral sets it
    mnrl_affinity_power: Optional[np.ndarray] = None
    mnrl_temkin: Optional[np.ndarray] = None
    # isotherm sorption (reaction%eqkd*, reaction_database.F90:3393-3450):
    # [D] arrays, empty without ISOTHERM_REACTIONS
    kd_specid: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    kd_type: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    kd_coef: np.ndarray = field(default_factory=lambda: np.zeros(0))
    kd_langmuir_b: np.ndarray = field(default_factory=lambda: np.zeros(0))
    kd_freundlich_n: np.ndarray = field(default_factory=lambda: np.zeros(0))
    kd_mineral: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    soil_particle_density: float = 0.0            # kg/m^3 (ROCK_DENSITY)
    # rate prefactors (mineral%kinmnrl_num_prefactors, kinmnrl_prefactor_id,
    # kinmnrl_pref_*): [K], [K][P][Q+1], [K][P][Q], [K][P]; None without any
    mnrl_num_prefactors: Optional[np.ndarray] = None
    mnrl_prefactor_id: Optional[np.ndarray] = None
    mnrl_pref_alpha: Optional[np.ndarray] = None
    mnrl_pref_beta: Optional[np.ndarray] = None
    mnrl_pref_atten: Optional[np.ndarray] = None
    mnrl_pref_rate: Optional[np.ndarray] = None
    mnrl_pref_activation_energy: Optional[np.ndarray] = None
    # MIN_SCALE_FACTOR: allocated for every kinetic mineral (1 for those
    # that do not set it) when any does (reaction_database.F90:1973-1983)
    mnrl_min_scale_factor: Optional[np.ndarray] = None
    # lanes per cell group in the per-cell kernels (0: the smallest that
    # holds N); a tuning knob of this library
    group_lanes: int = 0
    _desc: Optional[abi.ReactionDesc] = None

    @property
    def N(self) -> int:
        return len(self.primary_names)

    @property
    def M(self) -> int:
        return len(self.complex_names)

    @property
    def K(self) -> int:
        return len(self.mineral_names)

    @property
    def R(self) -> int:
        return self.ionx_CEC.shape[0]

    @property
    def D(self) -> int:
        return self.kd_specid.shape[0]

    @property
    def has_sorption(self) -> bool:
        """Sorbed totals are carried (reaction%neqsorb > 0): ion exchange or
        isotherms."""
        return self.R > 0 or self.D > 0

    def desc(self) -> abi.ReactionDesc:
        d = abi.ReactionDesc()
        d.naqcomp, d.neqcplx, d.nkinmnrl, d.neqionxrxn = self.N, self.M, self.K, self.R
        d.max_rxn_spec = self.cplx_stoich.shape[1]
        d.max_cation = self.ionx_k.shape[1]

        def dp(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_double))
            assert a.dtype == np.float64 and a.flags.c_contiguous
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def ip(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_int32))
            assert a.dtype == np.int32 and a.flags.c_contiguous
            return a.ctypes.data_as(C.POINTER(C.c_int32))

        d.primary_spec_Z, d.primary_spec_a0 = dp(self.Z), dp(self.a0)
        d.eqcplxspecid, d.eqcplxstoich = ip(self.cplx_specid), dp(self.cplx_stoich)
        d.eqcplxh2ostoich, d.eqcplx_Z, d.eqcplx_a0 = dp(self.cplx_h2ostoich), dp(self.cplx_Z), dp(self.cplx_a0)
        d.eqcplx_logK, d.eqcplx_logKcoef = dp(self.cplx_logK), dp(self.cplx_logKcoef)
        d.kinmnrlspecid, d.kinmnrlstoich = ip(self.mnrl_specid), dp(self.mnrl_stoich)
        d.kinmnrlh2ostoich, d.kinmnrl_logK = dp(self.mnrl_h2ostoich), dp(self.mnrl_logK)
        d.kinmnrl_logKcoef = dp(self.mnrl_logKcoef)
        d.kinmnrl_rate_constant, d.kinmnrl_molar_vol = dp(self.mnrl_rate_constant), dp(self.mnrl_molar_vol)
        d.kinmnrl_activation_energy = dp(self.mnrl_activation_energy)
        d.kinmnrl_affinity_threshold = dp(self.mnrl_affinity_threshold)
        d.kinmnrl_rate_limiter = dp(self.mnrl_rate_limiter)
        d.kinmnrl_affinity_power = dp(self.mnrl_affinity_power)
        d.kinmnrl_Temkin_const = dp(self.mnrl_temkin)
        d.eqionx_rxn_cationid, d.eqionx_rxn_k = ip(self.ionx_cationid), dp(self.ionx_k)
        d.eqionx_rxn_CEC, d.eqionx_rxn_to_surf = dp(self.ionx_CEC), ip(self.ionx_to_surf)
        d.debyeA, d.debyeB, d.debyeBdot = self.debye
        d.use_log_formulation = self.use_log_formulation
        d.act_coef_update_frequency = self.act_freq
        d.act_coef_update_algorithm = self.act_alg
        d.numerical_jacobian = self.numerical_jacobian
        d.max_dlnC = self.max_dlnC
        d.max_relative_change_tolerance = self.max_rel_change_tol
        d.max_residual_tolerance = self.max_residual_tol
        d.neqkdrxn = self.D
        if self.D:
            d.eqkdspecid, d.eqkdtype = ip(self.kd_specid), ip(self.kd_type)
            d.eqkddistcoef, d.eqkdlangmuirb = dp(self.kd_coef), dp(self.kd_langmuir_b)
            d.eqkdfreundlichn, d.eqkdmineral = dp(self.kd_freundlich_n), ip(self.kd_mineral)
        d.soil_particle_density = self.soil_particle_density
        if self.mnrl_num_prefactors is not None:
            d.max_prefactors = self.mnrl_pref_rate.shape[1]
            d.max_prefactor_species = self.mnrl_pref_alpha.shape[2]
            d.kinmnrl_num_prefactors, d.kinmnrl_prefactor_id = (ip(self.mnrl_num_prefactors),
                                                                ip(self.mnrl_prefactor_id))
            d.kinmnrl_pref_alpha, d.kinmnrl_pref_beta = dp(self.mnrl_pref_alpha), dp(self.mnrl_pref_beta)
            d.kinmnrl_pref_atten_coef, d.kinmnrl_pref_rate = dp(self.mnrl_pref_atten), dp(self.mnrl_pref_rate)
            d.kinmnrl_pref_activation_energy = dp(self.mnrl_pref_activation_energy)
        d.kinmnrl_min_scale_factor = dp(self.mnrl_min_scale_factor)
        d.group_lanes = int(self.group_lanes)
        self._desc = d
        return d

    def flops_per_iteration(self) -> Dict[str, float]:
        """Algorithmic work of one Newton iteration of one cell, counted from
        the realised tables (FMA = 2 flops; exp/log/pow counted separately)."""
        N, M, K = self.N, self.M, self.K
        s = self.cplx_specid[:, 0].astype(np.float64) if M else np.zeros(0)
        t = self.mnrl_specid[:, 0].astype(np.float64) if K else np.zeros(0)
        rtotal = float(np.sum(4 * s + 2 + 4 * s + 2 * s * s)) + N * N + 2 * N
        accum = N + N * N
        ionx = 0.0
        for r in range(self.R):
            c = float(self.ionx_cationid[r, 0])
            ionx += 10 * (c * 6) + 3 * c * c
        mnrl = float(np.sum(2 * t + 12 + 2 * t * t + 3 * t)) if K else 0.0
        rsolve = 2 * N * N + N + N * N + (2.0 / 3.0) * N ** 3 + N * N + 2 * N * N
        update = 3 * N
        flops = rtotal + accum + ionx + mnrl + rsolve + update
        transc = 2 * N + float(np.sum(1 + s)) + K + N
        return {"flops": flops, "transcendentals": transc}


def constraint_rxn(db: Database, name: str, primaries: Sequence[str], kind: str = "mineral",
                   temperature: float = 25.0):
    """abi.ConstraintRxn of a mineral or gas constraint (the species of the
    database reaction, water dropped; log K at the reference temperature as
    build_tables takes it, plus the Maier-Kelly coefficients)."""
    r = Basis(db, primaries).reaction(name, db.minerals if kind == "mineral" else db.gases)
    pidx = {n: i for i, n in enumerate(primaries)}
    x = abi.ConstraintRxn()
    k = 0
    for st, sp in r.species:
        if sp in H2O_NAMES:
            continue
        if sp not in pidx:
            raise ValueError(f"{name}: species {sp} is not a primary")
        x.specid[k] = pidx[sp]
        x.stoich[k] = st
        k += 1
    x.nspec = k
    temps = db.temperatures
    itemp = next((i for i, t in enumerate(temps) if abs(t - temperature) < 1e-10), None)
    coef = logk_fit(temps, r.logK)
    if itemp is not None and abs(r.logK[itemp] - 500.0) > 1e-10:
        x.logK = r.logK[itemp]
    else:
        x.logK = float(logk_eval(coef, temperature))
    x.has_logK_coef = 1
    for i in range(5):
        x.logK_coef[i] = coef[i]
    return x


def constraint_rxns(t: "ReactionTables", rxns: Dict[int, "abi.ConstraintRxn"]):
    """ctypes array [N] of constraint reactions (components without one hold
    an empty entry)."""
    arr = (abi.ConstraintRxn * t.N)()
    for j, x in rxns.items():
        arr[j] = x
    return arr


def build_tables(db: Database, primaries: Sequence[str], secondaries: Sequence[str] = (),
                 minerals: Dict[str, MineralKinetics] | None = None,
                 ion_exchange: Sequence[IonExchangeRxn] = (), *, isotherms: Sequence[KDRxn] = (),
                 soil_particle_density: float = 0.0, temperature: float = 25.0,
                 log_formulation: bool = True, act_freq: int = abi.PFX_ACT_OFF,
                 act_alg: int = abi.PFX_ACT_ALG_LAG, use_bdot: bool = True,
                 fit_logk: bool = False) -> ReactionTables:
    """Flatten a deck's CHEMISTRY block against a database, every reaction
    written in the deck's primaries (Basis).  log K at the
    reference temperature is the tabulated value when the temperature is a
    database temperature (ReactionInitializeLogK, reaction_aux.F90:1438-1450),
    otherwise the Maier-Kelly fit.  `isotherms` are the ISOTHERM_REACTIONS
    (their mineral names resolved among the kinetic minerals,
    GetKineticMineralIDFromName as reaction_database.F90:3434-3442 uses it);
    `soil_particle_density` is the ROCK_DENSITY they scale by."""
    minerals = dict(minerals or {})
    pidx = {n: i for i, n in enumerate(primaries)}
    basis = Basis(db, primaries)
    N, M, K, R = len(primaries), len(secondaries), len(minerals), len(ion_exchange)
    S = 1
    for n in secondaries:
        S = max(S, sum(1 for _, sp in basis.reaction(n, db.complexes).species if sp not in H2O_NAMES))
    for n in minerals:
        S = max(S, sum(1 for _, sp in basis.reaction(n, db.minerals).species if sp not in H2O_NAMES))
    temps = db.temperatures
    itemp = next((i for i, t in enumerate(temps) if abs(t - temperature) < 1e-10), None)

    def logk_at(lk):
        if itemp is not None and abs(lk[itemp] - 500.0) > 1e-10:
            return lk[itemp]
        return float(logk_eval(logk_fit(temps, lk), temperature))

    Z = np.array([basis.primary(n).Z for n in primaries], dtype=np.float64)
    a0 = np.array([basis.primary(n).a0 for n in primaries], dtype=np.float64)

    def rxn_arrays(names, src):
        spec = np.zeros((len(names), S + 1), dtype=np.int32)
        nu = np.zeros((len(names), S), dtype=np.float64)
        h2o = np.zeros(len(names), dtype=np.float64)
        logK = np.zeros(len(names), dtype=np.float64)
        coef = np.zeros((len(names), 5), dtype=np.float64)
        for i, n in enumerate(names):
            r = basis.reaction(n, src)
            k = 0
            for st, sp in r.species:
                if sp in H2O_NAMES:
                    h2o[i] = st
                    continue
                if sp not in pidx:
                    raise ValueError(f"{n}: species {sp} is not a primary")
                spec[i, 1 + k] = pidx[sp]
                nu[i, k] = st
                k += 1
            spec[i, 0] = k
            logK[i] = logk_at(r.logK)
            coef[i] = logk_fit(temps, r.logK)
        return spec, nu, h2o, logK, coef

    cs, cn, ch, cl, cc = rxn_arrays(list(secondaries), db.complexes)
    mnames = list(minerals)
    ms, mn, mh, ml, mc = rxn_arrays(mnames, db.minerals)
    Cmax = max([len(r.cations) for r in ion_exchange] + [1])
    ionx_id = np.zeros((R, Cmax + 1), dtype=np.int32)
    ionx_k = np.zeros((R, Cmax), dtype=np.float64)
    for r, rx in enumerate(ion_exchange):
        ionx_id[r, 0] = len(rx.cations)
        for j, (nm, k) in enumerate(rx.cations):
            ionx_id[r, 1 + j] = pidx[nm]
            ionx_k[r, j] = k
    mnames_ = list(minerals)
    for kr in isotherms:
        if kr.species not in pidx:
            raise ValueError(f"species {kr.species} in kd reaction not found among primary species list")
        if kr.mineral and kr.mineral not in mnames_:
            raise ValueError(f"mineral {kr.mineral} listed in kd reaction not found in the kinetic minerals")
    kd_kw = dict(
        kd_specid=np.array([pidx[kr.species] for kr in isotherms], dtype=np.int32),
        kd_type=np.array([kr.type for kr in isotherms], dtype=np.int32),
        kd_coef=np.array([kr.kd for kr in isotherms], dtype=np.float64),
        kd_langmuir_b=np.array([kr.langmuir_b for kr in isotherms], dtype=np.float64),
        kd_freundlich_n=np.array([kr.freundlich_n for kr in isotherms], dtype=np.float64),
        kd_mineral=np.array([mnames_.index(kr.mineral) if kr.mineral else -1 for kr in isotherms], dtype=np.int32),
        soil_particle_density=float(soil_particle_density))
    pref_kw = prefactor_arrays([minerals[n] for n in mnames], pidx)
    return ReactionTables(
        primary_names=list(primaries), complex_names=list(secondaries), mineral_names=mnames,
        Z=Z, a0=a0, cplx_specid=cs, cplx_stoich=cn, cplx_h2ostoich=ch,
        cplx_Z=np.array([basis.reaction(n, db.complexes).Z for n in secondaries], dtype=np.float64),
        cplx_a0=np.array([basis.reaction(n, db.complexes).a0 for n in secondaries], dtype=np.float64),
        cplx_logK=cl, cplx_logKcoef=cc if fit_logk else None,
        mnrl_specid=ms, mnrl_stoich=mn, mnrl_h2ostoich=mh, mnrl_logK=ml,
        mnrl_logKcoef=mc if fit_logk else None,
        mnrl_rate_constant=np.array([minerals[n].rate_constant for n in mnames], dtype=np.float64),
        mnrl_molar_vol=np.array([db.minerals[n].molar_volume * 1e-6 for n in mnames], dtype=np.float64),
        mnrl_activation_energy=np.array([minerals[n].activation_energy for n in mnames], dtype=np.float64),
        mnrl_affinity_threshold=np.array([minerals[n].affinity_threshold for n in mnames], dtype=np.float64),
        mnrl_rate_limiter=np.array([minerals[n].rate_limiter for n in mnames], dtype=np.float64),
        mnrl_affinity_power=_optional_per_mineral([minerals[n].affinity_power for n in mnames]),
        mnrl_temkin=_optional_per_mineral([minerals[n].temkin for n in mnames]),
        ionx_cationid=ionx_id, ionx_k=ionx_k,
        ionx_CEC=np.array([rx.CEC for rx in ion_exchange], dtype=np.float64),
        ionx_to_surf=np.array([mnames.index(rx.mineral) if rx.mineral else -1 for rx in ion_exchange],
                              dtype=np.int32),
        debye=debye_params(temperature, use_bdot), use_log_formulation=int(log_formulation),
        act_freq=act_freq, act_alg=act_alg,
        mnrl_min_scale_factor=_optional_per_mineral([minerals[n].min_scale_factor for n in mnames]),
        **kd_kw, **pref_kw)


def prefactor_arrays(kinetics: Sequence[MineralKinetics], pidx: Dict[str, int]) -> Dict[str, np.ndarray]:
    """The prefactor tables of the kinetic minerals (reaction_database.F90:
    1798-1829, 1944-1965, 2200-2240): counts, 0-based primary ids with the
    count in slot 0, alpha/beta/attenuation per species, rate constant and
    activation energy per prefactor, the mineral's own where the prefactor
    leaves them unset.  Empty when no mineral has a prefactor."""
    P = max([len(k.prefactors) for k in kinetics] + [0])
    if P == 0:
        return {}
    Q = max(len(p.species) for k in kinetics for p in k.prefactors)
    Q = max(Q, 1)
    K = len(kinetics)
    num = np.zeros(K, dtype=np.int32)
    pid = np.zeros((K, P, Q + 1), dtype=np.int32)
    alpha, beta, atten = (np.zeros((K, P, Q)) for _ in range(3))
    rate, ea = np.zeros((K, P)), np.zeros((K, P))
    for k, kin in enumerate(kinetics):
        num[k] = len(kin.prefactors)
        for p, pf in enumerate(kin.prefactors):
            rate[k, p] = kin.rate_constant if pf.rate is None else pf.rate
            ea[k, p] = kin.activation_energy if pf.activation_energy is None else pf.activation_energy
            pid[k, p, 0] = len(pf.species)
            for q, sp in enumerate(pf.species):
                if sp.name not in pidx:
                    raise ValueError(f"prefactor species {sp.name}: only primary species are supported")
                pid[k, p, 1 + q] = pidx[sp.name]
                alpha[k, p, q], beta[k, p, q], atten[k, p, q] = sp.alpha, sp.beta, sp.atten
    return dict(mnrl_num_prefactors=num, mnrl_prefactor_id=pid, mnrl_pref_alpha=alpha, mnrl_pref_beta=beta,
                mnrl_pref_atten=atten, mnrl_pref_rate=rate, mnrl_pref_activation_energy=ea)
