"""Global-implicit reactive transport of a deck's 1-D column on the device:
the reference's time loop (stepper.TimeStepper) around the device Newton
step with its time-step cut (pfx_gi_solver_timestep), with the per-step
updates the transport process model makes around the solve.

Per step, in the reference's order:
  PMRTInitializeTimestep (pm_rt.F90:361-415): the fixed accumulation of the
      old state (RTUpdateFixedAccumulation, reactive_transport.F90:726-848:
      the totals at the old free ions with the old activity coefficients),
      then, with ACTIVITY_COEFFICIENTS TIMESTEP, one activity update of every
      cell and every boundary state (RTUpdateActivityCoefficients,
      reactive_transport.F90:3620-3700) from its current complexes
  SNESSolve (timestepper_BE.F90:326): pfx_gi_solver_timestep; a boundary
      face carries the totals of its boundary state (the constraint's free
      ions with the boundary's own coefficients, RTUpdateAuxVars
      reactive_transport.F90:3850-4025); a zero-gradient outflow face
      carries the cell's own
  PMRTUpdateSolution (pm_rt.F90:1039-1040): the mineral volume fractions
      from the rates at the solution (RTUpdateKineticState)
  TimestepperBEUpdateDT: the next size from the step's Newton count

Set-up as CondControlAssignTranInitCond (condition_control.F90:725-950):
every cell takes the initial constraint's free ions with unit activity
coefficients, then two passes of activity update and aux-var refresh;
boundary states start from their constraint's free ions with unit
coefficients (reactive_transport_aux.F90 RTAuxVarInit) and their totals are
computed without activity (PMRTInitializeRun, pm_rt.F90:335-337).

The reference's transport linear solver on these columns is BCGS with
PETSc's serial default ILU(0) or a direct solve; on a 1-D block-tridiagonal
matrix ILU(0) has no fill and is the exact factorization, so both solve
each Newton system exactly.  The device solve (point-block Jacobi BiCGStab)
is run to a tight tolerance to stand for that.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi
from .deck import Deck


@dataclass
class ColumnResult:
    steps: int
    newton_iterations: int
    cuts: int
    history: List
    state: Dict[str, np.ndarray] = field(default_factory=dict)


def face_data(box, q_sign_west: float = 1.0):
    """Face fluxes and dispersion over distance of the column (TFluxCoef
    upwinding happens in the kernel; TDispersion / TDispersionBC here):
    internal faces D / dx (equal cells), a Dirichlet face D / (dx / 2), a
    dirichlet_zero_gradient face Dirichlet on inflow and zero-gradient on
    outflow, a zero-gradient face none; D = max(alpha_L |q| + s phi tau
    D_m, 1e-40).  Boundary fluxes are positive into the box."""
    n, q = box.nx, box.darcy
    D = max(box.dispersivity * abs(q) + 1.0 * box.porosity * box.tortuosity * box.diffusion, 1e-40)
    e = [np.zeros((1, 0, n)), np.zeros((0, 1, n))]
    qf = [np.full((1, 1, n - 1), q)] + e
    ddf = [np.full((1, 1, n - 1), D / box.dx)] + e
    bc_q, bc_dd, dirichlet = [], [], []
    for face, qin in (("west", q), ("east", -q)):
        bc = getattr(box, face)
        if bc is None:
            bc_q.append(None)
            bc_dd.append(None)
            dirichlet.append(False)
            continue
        typ = bc[0]
        dir_ = typ == "DIRICHLET" or (typ == "DIRICHLET_ZERO_GRADIENT" and qin >= 0.0)
        bc_q.append(np.full((1, 1), qin))
        bc_dd.append(np.full((1, 1), D / (0.5 * box.dx) if dir_ else 0.0))
        dirichlet.append(dir_)
    return qf, ddf, bc_q, bc_dd, dirichlet


class ColumnRun:
    """A deck's column on the device: tables, initial cells, boundary
    states, the grid and the solver handle."""

    def __init__(self, deck: Deck, db=None, newton: Optional[abi.NewtonOpts] = None, linear_rtol: float = 1e-13,
                 jacobian_flags: int = 0):
        import torch

        from . import runtime
        self.deck = deck
        self.db = db if db is not None else deck.database()
        self.t = deck.tables(self.db)
        self.box = deck.box()
        self.tab = runtime.Tables(self.t)
        self.den = deck.water_density()
        self.jacobian_flags = jacobian_flags
        box, t = self.box, self.t
        n = box.nx
        # every cell from the initial constraint (identical cells: one solve)
        one = self._speciate(box.initial)
        arr = {k: (None if v is None else np.repeat(v.cpu().numpy(), n, axis=0)) for k, v in one.t.items()}
        self.dc = runtime.DeviceCells.from_host(arr, n)
        self.dc.t["pri_act_coef"].fill_(1.0)
        if self.dc.t.get("sec_act_coef") is not None:
            self.dc.t["sec_act_coef"].fill_(1.0)
        if self.dc.t.get("ionx_ref_sorbed") is not None:
            self.dc.t["ionx_ref_sorbed"].fill_(1e-9)
        passes = 2 if t.act_freq != abi.PFX_ACT_OFF else 0
        runtime.auxvar(self.tab, self.dc, 0, self.opts(1.0))
        if passes:
            runtime.auxvar(self.tab, self.dc, passes, self.opts(1.0))
        # boundary states: constraint free ions, unit coefficients, totals
        # without activity
        qf, ddf, bc_q, bc_dd, dirichlet = face_data(box)
        self.bc_cells: List = [None, None]
        bc_total = [None, None]
        for i, face in enumerate(("west", "east")):
            bc = getattr(box, face)
            if bc is None or not dirichlet[i]:
                continue
            c = self._speciate(bc[1])
            c.t["pri_act_coef"].fill_(1.0)
            if c.t.get("sec_act_coef") is not None:
                c.t["sec_act_coef"].fill_(1.0)
            if c.t.get("ionx_ref_sorbed") is not None:
                c.t["ionx_ref_sorbed"].fill_(1e-9)
            runtime.auxvar(self.tab, c, 0, self.opts(1.0))
            self.bc_cells[i] = c
            bc_total[i] = c.t["total"][0].clone()
        self.grid = runtime.DeviceGrid(n, 1, 1, qf, ddf, area=(box.area, 1.0, 1.0), bc_q=bc_q, bc_dd=bc_dd,
                                       bc_total=bc_total, log_formulation=bool(t.use_log_formulation))
        self.solver = runtime.GiSolver(self.tab, self.grid)
        self.newton = newton if newton is not None else self.deck_newton_opts(deck)
        if self.newton.krylov.rtol <= 0.0:
            self.newton.krylov.rtol = linear_rtol
        if self.newton.krylov.max_iterations <= 0:
            self.newton.krylov.max_iterations = 2000
        # scratch totals for the activity update of the step start (the
        # fixed accumulation keeps the old state's totals)
        self._scratch = {k: torch.empty_like(self.dc.t[k]) for k in ("total", "total_sorb", "ionx_ref_sorbed")
                         if self.dc.t.get(k) is not None}

    @staticmethod
    def deck_newton_opts(deck: Deck) -> abi.NewtonOpts:
        """pfx_newton_opts of the deck's transport NEWTON_SOLVER card and
        ITOL_RELATIVE_UPDATE (zeros: PETSc's defaults, PFLOTRAN's tests off)."""
        nw = deck.newton_options()
        o = abi.NewtonOpts()
        o.atol, o.rtol, o.stol = nw.get("atol", 0.0), nw.get("rtol", 0.0), nw.get("stol", 0.0)
        o.max_iterations = int(nw.get("max_it", 0))
        o.itol_res = nw.get("itol_res", 0.0)
        o.itol_update = nw.get("itol_upd", 0.0)
        o.itol_rel_update = nw.get("rel_update", 0.0)
        if "max_norm" in nw:
            raise ValueError("NEWTON_SOLVER MAX_NORM: the device step keeps the reference's default 1e20")
        return o

    def opts(self, dt: float) -> abi.ReactOpts:
        b = self.box
        return abi.default_opts(dt=dt, porosity=b.porosity, saturation=1.0, volume=b.volume, den_kg=self.den,
                                jacobian_flags=self.jacobian_flags)

    def _speciate(self, name: str):
        from . import runtime
        setup = self.deck.constraint_setup(name, self.t, self.db)
        from .synth import host_state_arrays
        arr = host_state_arrays(self.t, 1, with_temperature=False)
        if self.t.K and setup.get("vf") is not None:
            arr["mnrl_volfrac"][0, :] = setup["vf"]
            arr["mnrl_area"][0, :] = setup["area"]
        o = self.opts(1.0)
        o.max_iterations = 10000   # the constraint Newton's own cap (reaction.F90:1941)
        if setup.get("guess") is not None:
            arr["guess"][0, :] = setup["guess"]
            o.guess_mode = abi.PFX_GUESS_ARRAY
        dc = runtime.DeviceCells.from_host(arr, 1)
        rc, st = runtime.equilibrate(self.tab, dc, setup["ctype"], np.asarray([setup["conc"]], dtype=np.float64), o,
                                     rxn=setup.get("rxn"), molal=setup.get("molal", 0))
        if rc != 0:
            raise runtime.PfxError(f"constraint {name} did not speciate: {st.as_dict()}")
        return dc

    def _activity_update(self):
        """One activity update of every cell and boundary state; the cells'
        totals (the fixed accumulation) and exchange state are kept."""
        from . import runtime
        view = dict(self.dc.t)
        for k, v in self._scratch.items():
            v.copy_(self.dc.t[k])
            view[k] = v
        runtime.auxvar(self.tab, runtime.DeviceCells(view, self.dc.ncells), 1, self.opts(1.0))
        for i, c in enumerate(self.bc_cells):
            if c is not None:
                runtime.auxvar(self.tab, c, 1, self.opts(1.0))
                self.grid.bc_total[i].copy_(c.t["total"][0])

    def run(self, max_steps: Optional[int] = None) -> ColumnResult:
        from . import runtime
        ts = self.deck.stepper()
        act = self.t.act_freq == abi.PFX_ACT_TIMESTEP
        its = cuts = 0
        while not ts.done:
            if max_steps is not None and ts.steps >= max_steps:
                break
            dt = ts.set_target_time()
            if act:
                self._activity_update()
            st = self.solver.timestep(self.dc, self.grid, self.opts(dt), self.newton)
            if st.reason <= 0:
                raise runtime.PfxError(f"step {ts.steps + 1} at t = {ts.time:.6e} s failed: {st.as_dict()}")
            ts.accept(st.dt, st.newton.iterations, st.cuts)
            its += st.newton_iterations
            cuts += st.cuts
            o = self.opts(st.dt)
            if self.t.K:
                runtime.update_kinetic_state(self.tab, self.dc, o)
            runtime.auxvar(self.tab, self.dc, 0, o)
        return ColumnResult(ts.steps, its, cuts, ts.history, self.dc.to_host())

    def close(self):
        self.solver.close()
