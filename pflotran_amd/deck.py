"""Reader of the chemistry part of a PFLOTRAN input deck.

A Fortran host hands the library tables that PFLOTRAN's own input pass has
already built (ReactionReadPass1/Pass2, src/pflotran/reaction.F90:113-1048;
TranConstraintRead, src/pflotran/transport_constraint.F90:163-685).  This
module gives Python callers the same route from a deck: it reads the cards
that fix the per-cell reaction system and its initial constraints, and turns
them into ``database.build_tables`` tables, ``pfx_equilibrate_batch``
constraint arrays and a time-step schedule.

Cards read (everything else in the deck is skipped line by line):

* ``CHEMISTRY``: PRIMARY_SPECIES, SECONDARY_SPECIES, GAS_SPECIES /
  PASSIVE_GAS_SPECIES / ACTIVE_GAS_SPECIES, MINERALS, MINERAL_KINETICS
  (RATE_CONSTANT [units], ACTIVATION_ENERGY [units], AFFINITY_THRESHOLD,
  RATE_LIMITER), SORPTION / ION_EXCHANGE_RXN (MINERAL, CEC, CATIONS with one
  REFERENCE cation of k = 1), SORPTION / ISOTHERM_REACTIONS (TYPE,
  DISTRIBUTION_COEFFICIENT [units], LANGMUIR_B, FREUNDLICH_N,
  KD_MINERAL_NAME), DATABASE, LOG_FORMULATION,
  ACTIVITY_COEFFICIENTS [OFF|LAG|NEWTON|TIMESTEP|NEWTON_ITERATION], NO_BDOT,
  MOLAL, MAX_DLNC, MAX_RELATIVE_CHANGE_TOLERANCE, MAX_RESIDUAL_TOLERANCE;
  OUTPUT and REDOX_SPECIES are skipped as the reference skips them.
* ``CONSTRAINT name``: CONCENTRATIONS (F/FREE, T/TOTAL, TOTAL_SORB, P/PH,
  L/LOG, M/MINERAL/MNRL name, G/GAS name, Z/CHG; TOTAL when the type is
  left out) and MINERALS (volume fraction, area [units]).
* ``MATERIAL_PROPERTY name``: POROSITY, ROCK_DENSITY.
* ``REFERENCE_DENSITY`` [units]: the water density of a run without flow.
* ``SUBSURFACE_TRANSPORT`` options: GLOBAL_IMPLICIT, NUMERICAL_JACOBIAN.
* ``TIME``: FINAL_TIME, INITIAL_TIMESTEP_SIZE, MAXIMUM_TIMESTEP_SIZE.
* ``REFERENCE_TEMPERATURE``.

Chemistry outside the per-cell path (surface complexation, general,
microbial and decay reactions, immobile species, colloids, sandboxes,
porosity updates, supercritical CO2 and dataset constraints, free-ion
guesses) raises ``DeckError`` naming the keyword rather than being dropped.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Callable, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .database import (Database, DbSpecies, IonExchangeRxn, KDRxn, MineralKinetics, Prefactor, PrefactorSpecies,
                       ReactionTables, build_tables,
                       constraint_rxn, constraint_rxns, read_database, water_density_ifc67)


class DeckError(ValueError):
    """A deck the reader cannot turn into the per-cell reaction system."""


# --------------------------------------------------------------------------
# units (units.F90:400-475, the SI factors of the categories used here)

_SI = {
    # volume -> m^3
    "cm^3": 1e-6, "ml": 1e-6, "mL": 1e-6, "l": 1e-3, "L": 1e-3, "dm^3": 1e-3, "m^3": 1.0,
    # area -> m^2
    "cm^2": 1e-4, "dm^2": 1e-2, "m^2": 1.0, "km^2": 1e6,
    # length -> m
    "km": 1e3, "m": 1.0, "met": 1.0, "meter": 1.0, "dm": 1e-1, "cm": 1e-2, "mm": 1e-3,
    # time -> s
    "s": 1.0, "sec": 1.0, "second": 1.0, "min": 60.0, "minute": 60.0,
    "h": 3600.0, "hr": 3600.0, "hour": 3600.0, "d": 86400.0, "day": 86400.0,
    "w": 7 * 86400.0, "week": 7 * 86400.0,
    "mo": 365.0 / 12.0 * 86400.0, "month": 365.0 / 12.0 * 86400.0,
    "y": 365.0 * 86400.0, "yr": 365.0 * 86400.0, "year": 365.0 * 86400.0,
    # energy -> J
    "J": 1.0, "kJ": 1e3, "MJ": 1e6,
    # amount -> mol
    "mol": 1.0, "mole": 1.0, "moles": 1.0, "kmol": 1e3,
    # mass -> kg
    "ug": 1e-9, "mg": 1e-6, "g": 1e-3, "kg": 1.0,
}


def _part_si(part: str, unit: str) -> float:
    """SI factor of one side of a unit, its dash-joined units multiplied
    (UnitsConvert, units.F90:216-330)."""
    f = 1.0
    for tok in part.split("-"):
        if tok not in _SI:
            raise DeckError(f"unit '{tok}' of '{unit}' is not a unit this reader converts")
        f = f * _SI[tok]
    return f


def units_factor(unit: str, internal: str) -> float:
    """Factor taking a value in `unit` to `internal` (UnitsConvertToInternal
    and UnitsConvertParse, units.F90:17-160): the numerators' ratio over the
    denominators' ratio, each ratio user SI over internal SI."""
    u, i = unit.split("/"), internal.split("/")
    if len(u) > 2 or len(u) != len(i):
        raise DeckError(f"unit '{unit}' does not have the form of '{internal}'")
    num = _part_si(u[0], unit) / _part_si(i[0], internal)
    den = _part_si(u[1], unit) / _part_si(i[1], internal) if len(u) == 2 else 1.0
    return num / den


# --------------------------------------------------------------------------
# lines


def _num(tok: str) -> float:
    try:
        return float(tok.lower().replace("d", "e"))
    except ValueError:
        raise DeckError(f"'{tok}' is not a number") from None


def _lines(text: str, base_dir: Optional[str] = None, depth: int = 0) -> List[List[str]]:
    """Token lists of the deck's lines as InputReadPflotranString yields
    them (input_aux.F90:642-732): lines opening with '#' or '!' and blank
    lines skipped, a line cut at its first '#' or '!', SKIP ... NOSKIP
    regions (nested) dropped, an EXTERNAL_FILE line replaced by that file's
    lines (relative to the deck's folder)."""
    out, skip = [], 0
    for raw in text.splitlines():
        s = raw.strip()
        if not s or s[0] in "#!":
            continue
        for c in "#!":
            i = s.find(c)
            if i >= 0:
                s = s[:i]
        toks = s.replace("\t", " ").split()
        if not toks:
            continue
        w = toks[0].upper()
        if w.startswith("SKIP"):
            skip += 1
            continue
        if skip:
            if w.startswith("NOSK"):
                skip -= 1
            continue
        if w.startswith("NOSK"):
            continue
        if w.startswith("EXTERNAL_FILE"):
            if base_dir is None or len(toks) < 2:
                raise DeckError("EXTERNAL_FILE needs the deck's path (read_deck)")
            if depth > 8:
                raise DeckError("EXTERNAL_FILE nested too deeply")
            sub = os.path.join(base_dir, toks[1])
            with open(sub) as f:
                out += _lines(f.read(), os.path.dirname(os.path.abspath(sub)), depth + 1)
            continue
        out.append(toks)
    return out


def _is_exit(toks: List[str]) -> bool:
    """InputCheckExit (input_aux.F90:1236-1273): '/', END or END_*."""
    w = toks[0]
    return w.startswith("/") or w.upper() == "END" or w.upper().startswith("END_")


class _Cursor:
    def __init__(self, lines: List[List[str]]):
        self.lines, self.i = lines, 0

    def next(self, where: str) -> List[str]:
        if self.i >= len(self.lines):
            raise DeckError(f"end of deck inside {where}")
        t = self.lines[self.i]
        self.i += 1
        return t

    def block(self, where: str) -> Iterator[List[str]]:
        """Lines of a block up to its exit line (consumed)."""
        while True:
            t = self.next(where)
            if _is_exit(t):
                return
            yield t

    def skip_block(self, where: str) -> None:
        for _ in self.block(where):
            pass


# --------------------------------------------------------------------------
# deck contents


@dataclass
class SpeciesConstraint:
    name: str
    value: float
    ctype: int
    aux: Optional[str] = None      # mineral or gas of an M/G constraint


@dataclass
class MineralConstraint:
    name: str
    volume_fraction: float
    area: float                    # m^2/m^3 bulk


@dataclass
class Constraint:
    name: str
    concentrations: List[SpeciesConstraint] = field(default_factory=list)
    minerals: List[MineralConstraint] = field(default_factory=list)
    free_ion_guess: Dict[str, float] = field(default_factory=dict)


@dataclass
class Chemistry:
    primaries: List[str] = field(default_factory=list)
    secondaries: List[str] = field(default_factory=list)
    gases: List[str] = field(default_factory=list)
    minerals: List[str] = field(default_factory=list)
    kinetics: Dict[str, MineralKinetics] = field(default_factory=dict)
    ion_exchange: List[IonExchangeRxn] = field(default_factory=list)
    isotherms: List[KDRxn] = field(default_factory=list)
    database: Optional[str] = None
    log_formulation: bool = False                 # reaction_aux.F90:433
    act_freq: int = abi.PFX_ACT_OFF               # reaction_aux.F90:418
    act_alg: int = abi.PFX_ACT_ALG_LAG            # reaction_aux.F90:419
    use_bdot: bool = True                         # reaction_aux.F90:417
    molal: bool = False
    max_dlnC: float = 5.0                         # reaction_aux.F90:569
    max_rel_change_tol: float = 1e-6
    max_residual_tol: float = 1e-12


@dataclass
class TimeCard:
    final: Optional[float] = None                 # s
    dt_init: Optional[float] = None
    dt_max: Optional[float] = None
    waypoint_dt: bool = False                     # MAXIMUM_TIMESTEP_SIZE ... at <time> lines
    # TIMESTEPPER TRANSPORT: TS_ACCELERATION and DT_FACTOR, with the
    # reference's defaults (timestepper_BE.F90:124-133)
    iaccel: int = 5
    tfac: Tuple[float, ...] = (2.0, 2.0, 2.0, 2.0, 2.0, 1.8, 1.6, 1.4, 1.2, 1.0, 1.0, 1.0, 1.0)


@dataclass
class RegionCard:
    name: str
    face: Optional[str] = None                    # WEST, EAST, SOUTH, NORTH, BOTTOM, TOP
    coordinates: Optional[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = None


@dataclass
class TransportConditionCard:
    name: str
    type: str                                     # DIRICHLET, DIRICHLET_ZERO_GRADIENT, ZERO_GRADIENT, ...
    constraint: str


@dataclass
class TransportCard:
    """The structured box and its transport: GRID, UNIFORM_VELOCITY,
    FLUID_PROPERTY, MATERIAL_PROPERTY tortuosity and dispersivity, REGION,
    TRANSPORT_CONDITION, INITIAL_CONDITION, BOUNDARY_CONDITION, STRATA and
    the OUTPUT times that make waypoints."""
    nxyz: Optional[Tuple[int, int, int]] = None
    bounds: Optional[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = None
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)   # m/s Darcy (UNIFORM_VELOCITY)
    diffusion: float = 1e-9                       # m^2/s, fluid.F90:57
    tortuosity: Dict[str, float] = field(default_factory=dict)      # material -> (default 1)
    dispersivity: Dict[str, float] = field(default_factory=dict)    # material -> longitudinal (m)
    regions: Dict[str, RegionCard] = field(default_factory=dict)
    conditions: Dict[str, TransportConditionCard] = field(default_factory=dict)
    initial: List[Tuple[str, str]] = field(default_factory=list)    # (transport condition, region)
    boundaries: List[Tuple[str, str]] = field(default_factory=list)
    strata: List[Tuple[str, str]] = field(default_factory=list)     # (region, material)
    output_times: List[float] = field(default_factory=list)         # s
    output_periods: List[float] = field(default_factory=list)       # s


@dataclass
class BoxSetup:
    """A 1-D structured box along x ready for the global-implicit run."""
    nx: int
    dx: float
    area: float                                   # face area dy dz (m^2)
    porosity: float
    tortuosity: float
    dispersivity: float
    diffusion: float
    darcy: float                                  # m/s along +x
    initial: str                                  # constraint of every cell
    west: Optional[Tuple[str, str]] = None        # (condition type, constraint) of the x-min face
    east: Optional[Tuple[str, str]] = None

    @property
    def volume(self) -> float:
        return self.dx * self.area


@dataclass
class Deck:
    path: Optional[str]
    chemistry: Chemistry
    constraints: Dict[str, Constraint]
    porosity: Dict[str, float]
    time: TimeCard
    temperature: float = 25.0
    numerical_jacobian: bool = False
    transport: TransportCard = field(default_factory=TransportCard)
    reference_density: Optional[float] = None     # kg/m^3, REFERENCE_DENSITY
    rock_density: Dict[str, float] = field(default_factory=dict)   # material -> kg/m^3
    # transport NEWTON_SOLVER tolerances and ITOL_RELATIVE_UPDATE
    newton: Dict[str, float] = field(default_factory=dict)

    # ---------------------------------------------------------------- builds

    def newton_options(self) -> Dict[str, float]:
        """Keyword arguments of the global-implicit Newton (PETSc's defaults
        rtol 1e-8, stol 1e-8, atol 1e-50, 50 iterations where the deck sets
        none; ITOL, ITOL_UPDATE and ITOL_RELATIVE_UPDATE only where set)."""
        return dict(self.newton)

    def tracer_only(self) -> bool:
        """No complexes, minerals, sorption or gases: the reference does not
        switch on its full geochemistry (reaction.F90:899-906)."""
        c = self.chemistry
        return not (c.secondaries or c.minerals or c.kinetics or c.ion_exchange or c.isotherms or c.gases)

    def water_density(self) -> float:
        """den_kg of a run without flow: REFERENCE_DENSITY, else the water
        density at the reference temperature and pressure
        (factory_subsurface.F90:995-1000, init_subsurface_transport.F90:50)."""
        if self.reference_density is not None:
            return self.reference_density
        return water_density_ifc67(self.temperature)

    def soil_particle_density(self) -> float:
        """ROCK_DENSITY of the deck's material (material.F90:1371-1374), which
        scales the isotherms tied to a mineral; 0 when none is needed."""
        if not any(k.mineral for k in self.chemistry.isotherms):
            return 0.0
        mats = {m for _, m in self.transport.strata if m} or set(self.porosity)
        vals = {self.rock_density[m] for m in mats if m in self.rock_density}
        if len(vals) != 1 or len(mats) != 1:
            raise DeckError("an isotherm scaled by a mineral needs one material with a ROCK_DENSITY")
        return vals.pop()

    def database(self, path: Optional[str] = None) -> Database:
        """The thermodynamic database the deck names, resolved against the
        deck's folder (or `path` when given).  A tracer-only deck may name
        none: its primaries are uncharged and take part in no reaction."""
        p = path or self.chemistry.database
        if p is None:
            if not self.tracer_only():
                raise DeckError("CHEMISTRY has no DATABASE")
            return Database(temperatures=[25.0],
                            primary={n: DbSpecies(n, 0.0, 0.0, 0.0) for n in self.chemistry.primaries})
        if not os.path.isabs(p) and self.path:
            p = os.path.join(os.path.dirname(os.path.abspath(self.path)), p)
        return read_database(p)

    def tables(self, db: Database, fit_logk: bool = False) -> ReactionTables:
        c = self.chemistry
        kin = {n: c.kinetics[n] for n in self.kinetic_minerals()}
        # without a database the activity update is off (reaction.F90:920-921)
        act_freq = c.act_freq if c.database is not None else abi.PFX_ACT_OFF
        t = build_tables(db, c.primaries, c.secondaries, kin, c.ion_exchange, isotherms=c.isotherms,
                         soil_particle_density=self.soil_particle_density(), temperature=self.temperature,
                         log_formulation=c.log_formulation, act_freq=act_freq, act_alg=c.act_alg,
                         use_bdot=c.use_bdot, fit_logk=fit_logk)
        t.max_dlnC = c.max_dlnC
        t.max_rel_change_tol = c.max_rel_change_tol
        t.max_residual_tol = c.max_residual_tol
        t.numerical_jacobian = int(self.numerical_jacobian)
        return t

    def kinetic_minerals(self) -> List[str]:
        """Kinetic minerals in MINERALS-list order (the basis numbers
        kinetic minerals as it walks the mineral list)."""
        c = self.chemistry
        missing = [n for n in c.kinetics if n not in c.minerals]
        if missing:
            raise DeckError(f"MINERAL_KINETICS names minerals missing from MINERALS: {missing}")
        return [n for n in c.minerals if n in c.kinetics]

    def constraint_setup(self, name: str, tables: ReactionTables, db: Database) -> Dict:
        """Arrays of one constraint in primary-species order: ``ctype``,
        ``conc``, ``rxn`` (ctypes array of mineral/gas reactions, or None),
        ``vf`` and ``area`` per kinetic mineral (None when the constraint
        has no MINERALS block), ``molal``."""
        if name not in self.constraints:
            raise DeckError(f"no CONSTRAINT {name}")
        con = self.constraints[name]
        c = self.chemistry
        by_name = {s.name: s for s in con.concentrations}
        extra = [n for n in by_name if n not in c.primaries]
        if extra:
            raise DeckError(f"CONSTRAINT {name}: {extra} are not primary species")
        missing = [n for n in c.primaries if n not in by_name]
        if missing:
            raise DeckError(f"CONSTRAINT {name}: no concentration for {missing}")
        ctype, conc, rxns = [], [], {}
        for j, n in enumerate(c.primaries):
            s = by_name[n]
            ctype.append(s.ctype)
            conc.append(s.value)
            if s.ctype in (abi.PFX_CONSTRAINT_MINERAL, abi.PFX_CONSTRAINT_GAS):
                kind = "mineral" if s.ctype == abi.PFX_CONSTRAINT_MINERAL else "gas"
                rxns[j] = constraint_rxn(db, s.aux, c.primaries, kind, self.temperature)
        kin = self.kinetic_minerals()
        vf = area = None
        if con.minerals:
            # a MINERALS block must cover every kinetic mineral
            # (transport_constraint.F90:525-534); a constraint without one
            # (a boundary's) leaves the minerals alone
            vf, area = np.zeros(len(kin)), np.zeros(len(kin))
            got = {m.name: m for m in con.minerals}
            lack = [n for n in kin if n not in got]
            if lack:
                raise DeckError(f"CONSTRAINT {name}: MINERALS gives no volume fraction/area for {lack}")
            for k, n in enumerate(kin):
                vf[k], area[k] = got[n].volume_fraction, got[n].area
        guess = None
        if con.free_ion_guess:
            # every primary must be guessed (constraint.F90 guess check,
            # reaction.F90:1265-1288)
            lack = [n for n in c.primaries if n not in con.free_ion_guess]
            if lack:
                raise DeckError(f"CONSTRAINT {name}: FREE_ION_GUESS gives no value for {lack}")
            guess = np.array([con.free_ion_guess[n] for n in c.primaries], dtype=np.float64)
        return dict(ctype=ctype, conc=conc, rxn=constraint_rxns(tables, rxns) if rxns else None,
                    vf=vf, area=area, molal=int(c.molal), guess=guess)

    def next_dt(self, dt: float, newton_iterations: int) -> float:
        """The step size after a step of size `dt` that converged in
        `newton_iterations` Newton iterations and was not cut
        (PMRTUpdateTimestep without the volume-fraction governor,
        pm_rt.F90:631-647, called from TimestepperBEUpdateDT,
        timestepper_BE.F90:238-246): tfac(n) dt for n up to TS_ACCELERATION
        and the length of tfac, half of dt beyond, never more than 2 dt or
        the maximum step.  TS_ACCELERATION 0 keeps dt.  A count below 1 is
        taken as 1 (the reference would index tfac(0))."""
        tc = self.time
        if tc.iaccel == 0:
            return dt
        dt_max = tc.dt_max if tc.dt_max is not None else tc.dt_init
        n = max(int(newton_iterations), 1)
        dtt = tc.tfac[n - 1] * dt if (n <= tc.iaccel and n <= len(tc.tfac)) else 0.5 * dt
        return max(min(dtt, 2.0 * dt, dt_max), 1e-20)

    def waypoints(self):
        """Waypoints of the run (stepper.build_waypoints): the TIME card's
        maximum step, the final time and the OUTPUT times."""
        from .stepper import build_waypoints
        tc = self.time
        if tc.final is None or tc.dt_init is None:
            raise DeckError("TIME needs FINAL_TIME and INITIAL_TIMESTEP_SIZE")
        if tc.waypoint_dt:
            raise DeckError("MAXIMUM_TIMESTEP_SIZE ... at <time> waypoints are not modelled")
        dt_max = tc.dt_max if tc.dt_max is not None else tc.dt_init
        tr = self.transport
        wps = build_waypoints(tc.final, dt_max, tr.output_times, tr.output_periods[0] if tr.output_periods else None)
        for per in tr.output_periods[1:]:
            more = build_waypoints(tc.final, dt_max, (), per)
            merged = {w.time: w for w in wps}
            for w in more:
                if w.time in merged:
                    merged[w.time].force = merged[w.time].force or w.force
                else:
                    merged[w.time] = w
            wps = [merged[k] for k in sorted(merged)]
        return wps

    def stepper(self):
        """The transport time-step controller of the run (stepper.TimeStepper)."""
        from .stepper import TimeStepper
        tc = self.time
        wps = self.waypoints()
        dt_max = tc.dt_max if tc.dt_max is not None else tc.dt_init
        return TimeStepper(tc.final, tc.dt_init, dt_max, wps, iaccel=tc.iaccel, tfac=tuple(tc.tfac))

    def time_steps(self, newton_iterations: Optional[Callable[[int, float], int]] = None) -> List[float]:
        """Step sizes from 0 to FINAL_TIME for a run whose steps are never
        cut (stepper.TimeStepper: waypoints, the final time, growth by
        `next_dt` from each step's Newton count, which
        `newton_iterations(step_index, dt)` returns; 1 when not given)."""
        ts = self.stepper()
        steps = []
        while not ts.done:
            dt = ts.set_target_time()
            n = 1 if newton_iterations is None else newton_iterations(len(steps), dt)
            steps.append(dt)
            ts.accept(dt, n, 0)
        return steps

    def box(self) -> BoxSetup:
        """The deck's structured box as a 1-D column along x: one material,
        one initial condition over the whole box, boundary conditions on the
        WEST / EAST faces only (the faces pfx_grid carries), a uniform
        velocity along x."""
        tr = self.transport
        if tr.nxyz is None or tr.bounds is None:
            raise DeckError("GRID with NXYZ and BOUNDS (or DXYZ) expected")
        nx, ny, nz = tr.nxyz
        if ny != 1 or nz != 1:
            raise DeckError("only columns along x (NXYZ n 1 1) are set up")
        (x0, y0, z0), (x1, y1, z1) = tr.bounds
        if abs(tr.velocity[1]) > 0.0 or abs(tr.velocity[2]) > 0.0:
            raise DeckError("UNIFORM_VELOCITY across the column is not modelled")
        mats = {m for _, m in tr.strata} or set(self.porosity)
        if len(mats) != 1:
            raise DeckError("one material expected")
        mat = next(iter(mats))
        if len(tr.initial) != 1:
            raise DeckError("one INITIAL_CONDITION expected")
        ic = tr.conditions[tr.initial[0][0]].constraint
        faces = {}
        for cname, rname in tr.boundaries:
            reg = tr.regions.get(rname)
            if reg is None or reg.face not in ("WEST", "EAST"):
                raise DeckError(f"BOUNDARY_CONDITION on region {rname}: only WEST / EAST faces are modelled")
            cond = tr.conditions[cname]
            if cond.type not in ("DIRICHLET", "DIRICHLET_ZERO_GRADIENT", "ZERO_GRADIENT"):
                raise DeckError(f"TRANSPORT_CONDITION {cname}: TYPE {cond.type} is not modelled")
            faces[reg.face] = (cond.type, cond.constraint)
        return BoxSetup(nx=nx, dx=(x1 - x0) / nx, area=(y1 - y0) * (z1 - z0), porosity=self.porosity[mat],
                        tortuosity=tr.tortuosity.get(mat, 1.0), dispersivity=tr.dispersivity.get(mat, 0.0),
                        diffusion=tr.diffusion, darcy=tr.velocity[0], initial=ic, west=faces.get("WEST"),
                        east=faces.get("EAST"))


# --------------------------------------------------------------------------
# parser

_CONSTRAINT_TYPES = {
    "F": abi.PFX_CONSTRAINT_FREE, "FREE": abi.PFX_CONSTRAINT_FREE,
    "T": abi.PFX_CONSTRAINT_TOTAL, "TOTAL": abi.PFX_CONSTRAINT_TOTAL,
    "TOTAL_SORB": abi.PFX_CONSTRAINT_TOTAL_SORB,
    "P": abi.PFX_CONSTRAINT_PH, "PH": abi.PFX_CONSTRAINT_PH,
    "L": abi.PFX_CONSTRAINT_LOG, "LOG": abi.PFX_CONSTRAINT_LOG,
    "M": abi.PFX_CONSTRAINT_MINERAL, "MINERAL": abi.PFX_CONSTRAINT_MINERAL, "MNRL": abi.PFX_CONSTRAINT_MINERAL,
    "G": abi.PFX_CONSTRAINT_GAS, "GAS": abi.PFX_CONSTRAINT_GAS,
    "Z": abi.PFX_CONSTRAINT_CHARGE, "CHG": abi.PFX_CONSTRAINT_CHARGE,
}

# CHEMISTRY keywords of chemistry outside the per-cell path
_CHEM_UNSUPPORTED = {
    "GENERAL_REACTION", "MICROBIAL_REACTION", "RADIOACTIVE_DECAY_REACTION", "IMMOBILE_SPECIES",
    "COLLOIDS", "REACTION_SANDBOX", "CLM_REACTION", "SOLID_SOLUTIONS", "UPDATE_POROSITY",
    "UPDATE_TORTUOSITY", "UPDATE_PERMEABILITY", "ACTIVITY_H2O",
    "ACTIVITY_WATER", "TRUNCATE_CONCENTRATION", "OPERATOR_SPLIT", "OPERATOR_SPLITTING", "GEOTHERMAL_HPT",
    "UPDATE_MNRL_SURF_AREA_WITH_POR", "UPDATE_ARMOR_MINERAL_SURFACE", "UPDATE_ARMOR_MINERAL_SURFACE_FLAG",
    "EXPLICIT_ADVECTION", "DECOUPLED_EQUILIBRIUM_MINERALS", "USE_FULL_GEOCHEMISTRY", "MINIMUM_POROSITY",
}
# keywords with no effect on the per-cell system
# (UPDATE_MINERAL_SURFACE_AREA: a transport-only run updates the surface
# area only along with the porosity, pm_rt.F90:238-247, 515-523, and
# UPDATE_POROSITY is refused above)
_CHEM_IGNORED = {"NO_CHECKPOINT_ACT_COEFS", "NO_RESTART_MINERAL_VOL_FRAC", "NO_CHECK_UPDATE",
                 "UPDATE_MINERAL_SURFACE_AREA"}


def _names(cur: _Cursor, where: str) -> List[str]:
    return [t[0] for t in cur.block(where)]


def _value_units(toks: List[str], i: int, internal: str, where: str) -> float:
    if len(toks) <= i:
        raise DeckError(f"{where}: value missing")
    v = _num(toks[i])
    if len(toks) > i + 1:
        v *= units_factor(toks[i + 1], internal)
    return v


def _mineral_kinetics(cur: _Cursor) -> Dict[str, MineralKinetics]:
    """MINERAL_KINETICS (reaction.F90:415-455 pass 1, reaction_mineral.F90
    MineralReadKinetics for the values)."""
    out = {}
    for t in cur.block("MINERAL_KINETICS"):
        name = t[0]
        k = MineralKinetics(rate_constant=float("nan"))
        for s in cur.block(f"MINERAL_KINETICS {name}"):
            w = s[0].upper()
            if w == "RATE_CONSTANT":
                rate = _num(s[1])
                if rate < 0.0:
                    rate = 10.0 ** rate
                if len(s) > 2:
                    rate *= units_factor(s[2], "mol/m^2-sec")
                k.rate_constant = rate
            elif w == "ACTIVATION_ENERGY":
                k.activation_energy = _value_units(s, 1, "J/mol", "ACTIVATION_ENERGY")
            elif w == "AFFINITY_THRESHOLD":
                k.affinity_threshold = _num(s[1])
            elif w == "RATE_LIMITER":
                k.rate_limiter = _num(s[1])
            elif w == "AFFINITY_POWER":
                k.affinity_power = _num(s[1])
            elif w == "TEMKIN_CONSTANT":
                k.temkin = _num(s[1])
            elif w == "MINERAL_SCALE_FACTOR":
                k.min_scale_factor = _num(s[1])
            elif w in ("SURFACE_AREA_VOL_FRAC_POWER", "SURFACE_AREA_POROSITY_POWER"):
                # used only where the surface area is updated, which a
                # transport-only run does only with UPDATE_POROSITY
                # (pm_rt.F90:238-247, 515-523), refused by this reader
                _num(s[1])
            elif w == "PREFACTOR":
                k.prefactors.append(_prefactor(cur, name))
            else:
                raise DeckError(f"MINERAL_KINETICS {name}: {w} is outside the per-cell path")
        # a prefactor without its own rate constant or activation energy
        # takes the mineral's (reaction_mineral.F90:343-360)
        for pf in k.prefactors:
            if pf.rate is None:
                if not np.isfinite(k.rate_constant):
                    raise DeckError(f"Both outer and inner prefactor rate constants uninitialized for kinetic "
                                    f"mineral {name}.")
                pf.rate = k.rate_constant
            if pf.activation_energy is None:
                pf.activation_energy = k.activation_energy
        if not np.isfinite(k.rate_constant):
            if not k.prefactors:
                raise DeckError(f"MINERAL_KINETICS {name}: no RATE_CONSTANT")
            k.rate_constant = 0.0   # unused: the prefactors carry the rate
        out[name] = k
    return out


def _prefactor(cur: _Cursor, mineral: str) -> Prefactor:
    """A PREFACTOR block (reaction_mineral.F90:227-335): RATE_CONSTANT
    (10^x when negative, units to mol/m^2-sec), ACTIVATION_ENERGY, and
    PREFACTOR_SPECIES blocks of ALPHA, BETA, ATTENUATION_COEF (default 0)."""
    pf = Prefactor(species=[])
    for t in cur.block(f"MINERAL_KINETICS {mineral} PREFACTOR"):
        w = t[0].upper()
        if w == "RATE_CONSTANT":
            rate = _num(t[1])
            if rate < 0.0:
                rate = 10.0 ** rate
            if len(t) > 2:
                rate *= units_factor(t[2], "mol/m^2-sec")
            pf.rate = rate
        elif w == "ACTIVATION_ENERGY":
            pf.activation_energy = _value_units(t, 1, "J/mol", "PREFACTOR ACTIVATION_ENERGY")
        elif w == "PREFACTOR_SPECIES":
            if len(t) < 2:
                raise DeckError("PREFACTOR_SPECIES: name missing")
            sp = PrefactorSpecies(t[1])
            for u in cur.block(f"PREFACTOR_SPECIES {t[1]}"):
                v = u[0].upper()
                if v == "ALPHA":
                    sp.alpha = _num(u[1])
                elif v == "BETA":
                    sp.beta = _num(u[1])
                elif v == "ATTENUATION_COEF":
                    sp.atten = _num(u[1])
                else:
                    raise DeckError(f"CHEMISTRY,MINERAL_KINETICS,PREFACTOR,SPECIES: {v} unknown")
            pf.species.append(sp)
        else:
            raise DeckError(f"CHEMISTRY,MINERAL_KINETICS,PREFACTOR: {w} unknown")
    return pf


def _ion_exchange(cur: _Cursor) -> IonExchangeRxn:
    """ION_EXCHANGE_RXN (reaction.F90:639-735): the REFERENCE cation is moved
    to the front and must have k = 1; CEC is read without units (eq/m^3)."""
    mineral, cec, cations, ref = None, None, [], None
    for t in cur.block("ION_EXCHANGE_RXN"):
        w = t[0].upper()
        if w == "MINERAL":
            mineral = t[1]
        elif w == "CEC":
            cec = _num(t[1])
        elif w == "CATIONS":
            for s in cur.block("CATIONS"):
                cations.append((s[0], _num(s[1])))
                if len(s) > 2:
                    if s[2].upper() != "REFERENCE":
                        raise DeckError(f"CATIONS: unknown word {s[2]}")
                    ref = s[0]
        else:
            raise DeckError(f"ION_EXCHANGE_RXN: {w} unknown")
    if ref is None:
        raise DeckError("Reference cation missing in Ion Exchange reaction.")
    i = next(i for i, (n, _) in enumerate(cations) if n == ref)
    if abs(cations[i][1] - 1.0) > 1e-40:
        raise DeckError(f'Reference cation "{ref}" must have k = 1.d0.')
    cations.insert(0, cations.pop(i))
    if cec is None:
        raise DeckError("ION_EXCHANGE_RXN without CEC")
    # on a mineral surface the site density is CEC times the mineral's
    # volume fraction (reaction.F90:4356-4363)
    return IonExchangeRxn(cations, cec, mineral)


def _isotherms(cur: _Cursor, units: List[str]) -> List[KDRxn]:
    """ISOTHERM_REACTIONS (reaction.F90:505-633).  As the reference reads
    them: every keyword line first resets the type to LINEAR, so the last
    line decides it (LANGMUIR_B and FREUNDLICH_N set theirs); units given
    after a DISTRIBUTION_COEFFICIENT stay in force for the reactions read
    after it (`units` carries them), converting K_D to kg/m^3 bulk, or to
    L/kg with a mineral."""
    out = []
    for t in cur.block("ISOTHERM_REACTIONS"):
        k = KDRxn(t[0], 0.0)
        for s in cur.block(f"ISOTHERM_REACTIONS {t[0]}"):
            w = s[0].upper()
            k.type = abi.PFX_SORPTION_LINEAR
            if w == "TYPE":
                v = s[1]
                types = {"LINEAR": abi.PFX_SORPTION_LINEAR, "LANGMUIR": abi.PFX_SORPTION_LANGMUIR,
                         "FREUNDLICH": abi.PFX_SORPTION_FREUNDLICH}
                if v not in types:    # compared as written
                    raise DeckError(f"CHEMISTRY,SORPTION,ISOTHERM_REACTIONS,TYPE: {v} unknown")
                k.type = types[v]
            elif w in ("DISTRIBUTION_COEFFICIENT", "KD"):
                k.kd = _num(s[1])
                if len(s) > 2:
                    units[0] = s[2]
            elif w == "LANGMUIR_B":
                k.langmuir_b = _num(s[1])
                k.type = abi.PFX_SORPTION_LANGMUIR
            elif w == "FREUNDLICH_N":
                k.freundlich_n = _num(s[1])
                k.type = abi.PFX_SORPTION_FREUNDLICH
            elif w == "KD_MINERAL_NAME":
                k.mineral = s[1]
            elif w in ("SEC_CONT_DISTRIBUTION_COEFFICIENT", "SEC_CONT_KD"):
                raise DeckError("Make sure MULTIPLE_CONTINUUM keyword is set, SECONDARY_CONTINUUM_KD.")
            else:
                raise DeckError(f"CHEMISTRY,SORPTION,ISOTHERM_REACTIONS: {w} unknown")
        if units[0]:
            k.kd *= units_factor(units[0], "L/kg" if k.mineral else "kg/m^3")
        # a one-letter mineral name is not looked up (reaction_database.F90:3434)
        if k.mineral is not None and len(k.mineral) < 2:
            k.mineral = None
        out.append(k)
    return out


def _chemistry(cur: _Cursor) -> Chemistry:
    c = Chemistry()
    kd_units = [""]
    for t in cur.block("CHEMISTRY"):
        w = t[0].upper()
        if w == "PRIMARY_SPECIES":
            c.primaries += _names(cur, w)
        elif w == "SECONDARY_SPECIES":
            c.secondaries += _names(cur, w)
        elif w in ("GAS_SPECIES", "PASSIVE_GAS_SPECIES", "ACTIVE_GAS_SPECIES"):
            c.gases += _names(cur, w)
        elif w == "MINERALS":
            c.minerals += _names(cur, w)
        elif w == "MINERAL_KINETICS":
            c.kinetics.update(_mineral_kinetics(cur))
        elif w == "SORPTION":
            for s in cur.block("SORPTION"):
                v = s[0].upper()
                if v == "ION_EXCHANGE_RXN":
                    c.ion_exchange.append(_ion_exchange(cur))
                elif v == "ISOTHERM_REACTIONS":
                    c.isotherms += _isotherms(cur, kd_units)
                elif v in ("SURFACE_COMPLEXATION_RXN", "DISTRIBUTION_COEFFICIENTS"):
                    raise DeckError(f"SORPTION {v} is outside the per-cell path")
                elif v in ("JUMPSTART_KINETIC_SORPTION", "NO_CHECKPOINT_KINETIC_SORPTION",
                           "NO_RESTART_KINETIC_SORPTION"):
                    pass
                else:
                    raise DeckError(f"SORPTION: {v} unknown")
        elif w == "DATABASE":
            c.database = t[1]
        elif w == "LOG_FORMULATION":
            c.log_formulation = True
        elif w == "ACTIVITY_COEFFICIENTS":
            # reaction.F90:768-789: the keyword alone means LAG at every step
            c.act_alg, c.act_freq = abi.PFX_ACT_ALG_LAG, abi.PFX_ACT_TIMESTEP
            for v in (x.upper() for x in t[1:]):
                if v == "OFF":
                    c.act_freq = abi.PFX_ACT_OFF
                elif v == "LAG":
                    c.act_alg = abi.PFX_ACT_ALG_LAG
                elif v == "NEWTON":
                    c.act_alg = abi.PFX_ACT_ALG_NEWTON
                elif v == "TIMESTEP":
                    c.act_freq = abi.PFX_ACT_TIMESTEP
                elif v == "NEWTON_ITERATION":
                    c.act_freq = abi.PFX_ACT_NEWTON_ITER
                else:
                    raise DeckError(f"CHEMISTRY,ACTIVITY_COEFFICIENTS: {v} unknown")
        elif w == "NO_BDOT":
            c.use_bdot = False
        elif w in ("MOLAL", "MOLALITY"):
            c.molal = True
        elif w == "MAX_DLNC":
            c.max_dlnC = _num(t[1])
        elif w in ("MAX_RELATIVE_CHANGE_TOLERANCE", "REACTION_TOLERANCE"):
            c.max_rel_change_tol = _num(t[1])
        elif w == "MAX_RESIDUAL_TOLERANCE":
            c.max_residual_tol = _num(t[1])
        elif w in ("OUTPUT", "REDOX_SPECIES"):
            cur.skip_block(w)
        elif w in _CHEM_IGNORED:
            pass
        elif w in _CHEM_UNSUPPORTED:
            raise DeckError(f"CHEMISTRY {w} is outside the per-cell path")
        else:
            raise DeckError(f"CHEMISTRY: keyword {w} unknown")
    return c


def _constraint(cur: _Cursor, name: str) -> Constraint:
    """CONSTRAINT (transport_constraint.F90:163-685)."""
    con = Constraint(name)
    for t in cur.block(f"CONSTRAINT {name}"):
        w = t[0].upper()
        if w in ("CONC", "CONCENTRATIONS"):
            for s in cur.block("CONCENTRATIONS"):
                if len(s) < 2:
                    raise DeckError(f"CONSTRAINT {name}: '{' '.join(s)}' lacks a value")
                ctype, aux = abi.PFX_CONSTRAINT_TOTAL, None
                if len(s) > 2:
                    v = s[2].upper()
                    if v == "S":
                        raise DeckError('"S" constraint type no longer supported as of March 4, 2013.')
                    if v not in _CONSTRAINT_TYPES:
                        raise DeckError(f"CONSTRAINT {name}: type {s[2]} is outside the per-cell path")
                    ctype = _CONSTRAINT_TYPES[v]
                    if ctype in (abi.PFX_CONSTRAINT_MINERAL, abi.PFX_CONSTRAINT_GAS):
                        if len(s) < 4:
                            raise DeckError(f"CONSTRAINT {name}: {s[0]} needs a mineral or gas name")
                        aux = s[3]
                    elif len(s) > 3 and s[3].upper() == "DATASET":
                        raise DeckError(f"CONSTRAINT {name}: DATASET concentrations are not read")
                con.concentrations.append(SpeciesConstraint(s[0], _num(s[1]), ctype, aux))
        elif w in ("MNRL", "MINERALS"):
            for s in cur.block("MINERALS"):
                if len(s) < 3 or "DATASET" in (x.upper() for x in s):
                    raise DeckError(f"CONSTRAINT {name}: mineral line '{' '.join(s)}' needs a volume fraction "
                                    "and an area (datasets are not read)")
                area = _num(s[2])
                if len(s) > 3:
                    area *= units_factor(s[3], "m^2/m^3")
                con.minerals.append(MineralConstraint(s[0], _num(s[1]), area))
        elif w == "FREE_ION_GUESS":
            # the constraint Newton's starting free ions (reaction.F90:1476-1482)
            for s in cur.block("FREE_ION_GUESS"):
                if len(s) < 2:
                    raise DeckError(f"CONSTRAINT {name}: guess '{' '.join(s)}' lacks a value")
                con.free_ion_guess[s[0]] = _num(s[1])
        else:
            raise DeckError(f"CONSTRAINT {name}: {w} is outside the per-cell path")
    return con


def _time(cur: _Cursor) -> TimeCard:
    tc = TimeCard()
    for t in cur.block("TIME"):
        w = t[0].upper()
        if w == "FINAL_TIME":
            tc.final = _value_units(t, 1, "s", w)
        elif w == "INITIAL_TIMESTEP_SIZE":
            tc.dt_init = _value_units(t, 1, "s", w)
        elif w == "MAXIMUM_TIMESTEP_SIZE":
            if len(t) > 3 and t[3].upper() == "AT":
                tc.waypoint_dt = True   # a later maximum from a waypoint on
            else:
                tc.dt_max = _value_units(t, 1, "s", w)
        elif w == "STEADY_STATE":
            raise DeckError("TIME STEADY_STATE is outside the per-cell path")
    return tc


def _xyz(toks: List[str], where: str) -> Tuple[float, float, float]:
    if len(toks) < 3:
        raise DeckError(f"{where}: three coordinates expected")
    return (_num(toks[0]), _num(toks[1]), _num(toks[2]))


def _grid(cur: _Cursor, tr: TransportCard) -> None:
    """GRID (grid.F90 GridRead): a structured box by NXYZ and BOUNDS (or a
    uniform DXYZ)."""
    dxyz = None
    for t in cur.block("GRID"):
        w = t[0].upper()
        if w == "TYPE":
            if len(t) < 2 or t[1].upper() != "STRUCTURED":
                raise DeckError("GRID: only TYPE structured boxes are read")
        elif w == "NXYZ":
            tr.nxyz = (int(_num(t[1])), int(_num(t[2])), int(_num(t[3])))
        elif w == "BOUNDS":
            lo = _xyz(cur.next("BOUNDS"), "BOUNDS")
            hi = _xyz(cur.next("BOUNDS"), "BOUNDS")
            cur.skip_block("BOUNDS")
            tr.bounds = (lo, hi)
        elif w == "DXYZ":
            vals = []
            for s_ in cur.block("DXYZ"):
                vals.append(s_)
            if len(vals) != 3 or any(len(v) != 1 for v in vals):
                raise DeckError("GRID DXYZ: only one uniform spacing per direction is read")
            dxyz = tuple(_num(v[0]) for v in vals)
        elif w == "ORIGIN":
            pass
        else:
            raise DeckError(f"GRID: {w} is not read")
    if dxyz is not None and tr.nxyz is not None:
        tr.bounds = ((0.0, 0.0, 0.0), tuple(n * d for n, d in zip(tr.nxyz, dxyz)))


def _region(cur: _Cursor, name: str) -> RegionCard:
    """REGION (region.F90 RegionRead): FACE and COORDINATES."""
    reg = RegionCard(name)
    for t in cur.block(f"REGION {name}"):
        w = t[0].upper()
        if w == "FACE":
            reg.face = t[1].upper()
        elif w == "COORDINATES":
            a = _xyz(cur.next("COORDINATES"), "COORDINATES")
            b = _xyz(cur.next("COORDINATES"), "COORDINATES")
            cur.skip_block("COORDINATES")
            reg.coordinates = (a, b)
        elif w == "COORDINATE":
            reg.coordinates = (_xyz(t[1:], "COORDINATE"),) * 2
        else:
            raise DeckError(f"REGION {name}: {w} is not read")
    return reg


def _transport_condition(cur: _Cursor, name: str) -> TransportConditionCard:
    """TRANSPORT_CONDITION (condition.F90 TranConditionRead): TYPE and the
    first entry of CONSTRAINT_LIST (a time-varying list is refused)."""
    typ, cons = None, []
    for t in cur.block(f"TRANSPORT_CONDITION {name}"):
        w = t[0].upper()
        if w == "TYPE":
            typ = t[1].upper()
        elif w == "CONSTRAINT_LIST":
            for s_ in cur.block("CONSTRAINT_LIST"):
                cons.append((_num(s_[0]), s_[1]))
        elif w == "CONSTRAINT":
            raise DeckError(f"TRANSPORT_CONDITION {name}: inline CONSTRAINT blocks are not read")
        elif w == "TIME_UNITS":
            pass
        else:
            raise DeckError(f"TRANSPORT_CONDITION {name}: {w} is not read")
    if typ is None or not cons:
        raise DeckError(f"TRANSPORT_CONDITION {name}: TYPE and CONSTRAINT_LIST expected")
    if len({c for _, c in cons}) > 1:
        raise DeckError(f"TRANSPORT_CONDITION {name}: constraints that change in time are not modelled")
    return TransportConditionCard(name, typ, cons[0][1])


def _coupler(cur: _Cursor, where: str) -> Tuple[str, str]:
    """INITIAL_CONDITION / BOUNDARY_CONDITION: (transport condition, region)."""
    cond = reg = None
    for t in cur.block(where):
        w = t[0].upper()
        if w == "TRANSPORT_CONDITION":
            cond = t[1]
        elif w == "REGION":
            reg = t[1]
        elif w == "FLOW_CONDITION":
            pass
        else:
            raise DeckError(f"{where}: {w} is not read")
    if cond is None or reg is None:
        raise DeckError(f"{where}: TRANSPORT_CONDITION and REGION expected")
    return cond, reg


def _output(cur: _Cursor, tr: TransportCard, depth: int = 0) -> None:
    """OUTPUT (output.F90:225-300 and factory_subsurface.F90:2270-2370): the
    times that become waypoints with output (TIMES, PERIODIC TIME,
    PERIODIC_OBSERVATION TIME), in the card or its SNAPSHOT_FILE,
    OBSERVATION_FILE and MASS_BALANCE_FILE blocks; the rest is output
    format and is not read."""
    for t in cur.block("OUTPUT"):
        w = t[0].upper()
        if w == "TIMES":
            f = units_factor(t[1], "s")
            vals = [_num(x) for x in t[2:]]
            tr.output_times += [v * f for v in vals]
        elif w in ("PERIODIC", "PERIODIC_OBSERVATION") and len(t) > 1 and t[1].upper() == "TIME":
            if len(t) > 4 and t[4].lower() == "between":
                raise DeckError("OUTPUT PERIODIC TIME ... between is not modelled")
            tr.output_periods.append(_num(t[2]) * units_factor(t[3], "s"))
        elif w in ("SNAPSHOT_FILE", "OBSERVATION_FILE", "MASS_BALANCE_FILE") and depth == 0:
            _output(cur, tr, depth + 1)
        elif w in ("VARIABLES", "AVERAGE_VARIABLES"):
            cur.skip_block(w)


def parse_deck(text: str, path: Optional[str] = None) -> Deck:
    cur = _Cursor(_lines(text, os.path.dirname(os.path.abspath(path)) if path else None))
    chem, constraints, porosity = None, {}, {}
    tc, temperature, numjac, stepper, newton = TimeCard(), 25.0, False, {}, {}
    tr = TransportCard()
    ref_den, rock = None, {}
    while cur.i < len(cur.lines):
        t = cur.next("deck")
        w = t[0].upper()
        if w == "CHEMISTRY":
            if chem is not None:
                raise DeckError("two CHEMISTRY cards")
            chem = _chemistry(cur)
        elif w == "CONSTRAINT" and len(t) > 1:
            constraints[t[1]] = _constraint(cur, t[1])
        elif w == "MATERIAL_PROPERTY" and len(t) > 1:
            name = t[1]
            depth = 0
            for s in cur.lines[cur.i:]:
                cur.i += 1
                v = s[0].upper()
                if _is_exit(s):
                    if depth == 0:
                        break
                    depth -= 1
                elif v in ("PERMEABILITY", "SECONDARY_CONTINUUM", "GEOMECHANICS_SUBSURFACE_PROPS"):
                    depth += 1
                elif v == "POROSITY" and depth == 0:
                    if len(s) > 2 and s[1].upper() == "DATASET":
                        raise DeckError(f"MATERIAL_PROPERTY {name}: POROSITY DATASET is not read")
                    porosity[name] = _num(s[1])
                elif v == "TORTUOSITY" and depth == 0:
                    tr.tortuosity[name] = _num(s[1])
                elif v == "ROCK_DENSITY" and depth == 0:
                    f = units_factor(s[2], "kg/m^3") if len(s) > 2 else 1.0
                    rock[name] = _num(s[1]) * f
                elif v == "LONGITUDINAL_DISPERSIVITY" and depth == 0:
                    tr.dispersivity[name] = _num(s[1])
        elif w == "SUBSURFACE_TRANSPORT":
            for s in cur.block("SUBSURFACE_TRANSPORT"):
                v = s[0].upper()
                if v == "NUMERICAL_JACOBIAN":
                    numjac = True
                elif v == "ITOL_RELATIVE_UPDATE":
                    newton["rel_update"] = _num(s[1])
                elif v == "ITOL_SCALED_RESIDUAL":
                    raise DeckError("SUBSURFACE_TRANSPORT ITOL_SCALED_RESIDUAL is not taken by the time loop")
                elif v in ("OPERATOR_SPLIT", "OPERATOR_SPLITTING"):
                    raise DeckError("operator-split transport is not reimplemented in the reference either")
        elif w == "TIME":
            tc = _time(cur)
        elif w == "NEWTON_SOLVER" and (len(t) < 2 or t[1].upper() == "TRANSPORT"):
            # solver.F90:841-900 (the SNES tolerances and PFLOTRAN's
            # infinity-norm tests of the transport Newton solve)
            keys = {"ATOL": "atol", "RTOL": "rtol", "STOL": "stol", "MAXIT": "max_it", "MAX_NORM": "max_norm",
                    "ITOL": "itol_res", "INF_TOL": "itol_res", "ITOL_RES": "itol_res", "INF_TOL_RES": "itol_res",
                    "ITOL_UPDATE": "itol_upd", "INF_TOL_UPDATE": "itol_upd"}
            for s in cur.block("NEWTON_SOLVER"):
                v = s[0].upper()
                if v in keys:
                    newton[keys[v]] = int(_num(s[1])) if v == "MAXIT" else _num(s[1])
                elif v == "DTOL":
                    pass   # SNES divergence of ||F|| beyond dtol ||F0||, not met on these decks
                elif v not in ("PRINT_DETAILED_CONVERGENCE", "PRINT_LINEAR_ITERATIONS", "NO_INFINITY_NORM",
                               "NO_PRINT_CONVERGENCE"):
                    raise DeckError(f"NEWTON_SOLVER: {v} is not taken by the time loop")
        elif w == "TIMESTEPPER" and (len(t) < 2 or t[1].upper() == "TRANSPORT"):
            for s in cur.block("TIMESTEPPER"):
                v = s[0].upper()
                if v == "TS_ACCELERATION":
                    stepper["iaccel"] = int(_num(s[1]))
                elif v == "DT_FACTOR":
                    if len(s) < 2:
                        raise DeckError("TIMESTEPPER DT_FACTOR: the factors go on the keyword's line")
                    stepper["tfac"] = tuple(_num(x) for x in s[1:])
        elif w == "REFERENCE_TEMPERATURE":
            temperature = _num(t[1])
        elif w == "REFERENCE_DENSITY":
            ref_den = _num(t[1]) * (units_factor(t[2], "kg/m^3") if len(t) > 2 else 1.0)
        elif w == "GRID":
            _grid(cur, tr)
        elif w == "UNIFORM_VELOCITY":
            v = _xyz(t[1:], "UNIFORM_VELOCITY")
            f = units_factor(t[4], "m/s") if len(t) > 4 else 1.0
            tr.velocity = tuple(x * f for x in v)
        elif w == "FLUID_PROPERTY":
            for s_ in cur.block("FLUID_PROPERTY"):
                if s_[0].upper() in ("DIFFUSION_COEFFICIENT", "LIQUID_DIFFUSION_COEFFICIENT"):
                    tr.diffusion = _num(s_[1])
        elif w == "REGION" and len(t) > 1:
            tr.regions[t[1]] = _region(cur, t[1])
        elif w == "TRANSPORT_CONDITION" and len(t) > 1:
            tr.conditions[t[1]] = _transport_condition(cur, t[1])
        elif w == "INITIAL_CONDITION":
            tr.initial.append(_coupler(cur, "INITIAL_CONDITION"))
        elif w == "BOUNDARY_CONDITION":
            tr.boundaries.append(_coupler(cur, "BOUNDARY_CONDITION"))
        elif w == "STRATA":
            reg = mat = None
            for s_ in cur.block("STRATA"):
                if s_[0].upper() == "REGION":
                    reg = s_[1]
                elif s_[0].upper() == "MATERIAL":
                    mat = s_[1]
            tr.strata.append((reg, mat))
        elif w == "OUTPUT":
            _output(cur, tr)
        elif w in ("OBSERVATION", "SOURCE_SINK", "INTEGRAL_FLUX"):
            cur.skip_block(w)   # their REGION lines are not region cards
    if chem is None:
        raise DeckError("deck has no CHEMISTRY card")
    for k, v in stepper.items():
        setattr(tc, k, v)
    return Deck(path, chem, constraints, porosity, tc, temperature, numjac, tr, ref_den, rock, newton)


def read_deck(path: str) -> Deck:
    with open(path) as f:
        return parse_deck(f.read(), path)
