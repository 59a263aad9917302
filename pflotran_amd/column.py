"""Operator-split reactive transport on a 1-D column: the transport half as an
explicit upwind step on the totals (numpy or torch), or implicit on the
device (pfx_os_transport_step, os_steps_device_implicit); the reaction half
as the batched RReact (pfx_react_batch) and the kinetic-state update
(pfx_update_kinetic_state).

The reference's OS loop is RTReact (reactive_transport.F90:1605-1806) after
its own transport solve; here the transport half is a plain explicit Euler
step of the global-implicit flux terms (TFluxCoef / TFlux,
transport.F90:368-439, 756-819): the totals move by
    (phi s V 1000) dT/dt = - r_flux(T),
with upwind advection, harmonic-mean dispersion between cells, a Dirichlet
inflow face (dispersion over half a cell) and a zero-gradient outflow face.
It is written once for numpy and torch arrays alike, so a CPU run and a
device run transport identically; it is not part of the measured kernels.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Column:
    ncells: int = 100
    dx: float = 1.0                # m
    area: float = 1.0              # m^2
    porosity: float = 0.25
    saturation: float = 1.0
    darcy: float = 3.1688e-8       # m/s (1 m/yr)
    diffusion: float = 1e-9        # m^2/s
    tortuosity: float = 1.0
    dispersivity: float = 0.0      # m

    @property
    def volume(self):
        return self.dx * self.area

    def coefficients(self):
        """(T_up, T_dn) of the internal faces and of the inflow/outflow
        faces, L/s (TFluxCoef with upwinding, q > 0 toward +x)."""
        q = self.darcy
        D = max(self.dispersivity * abs(q) + self.saturation * self.porosity * self.tortuosity * self.diffusion,
                1e-40)
        dd = D / self.dx                       # harmonic mean of equal sides over dx
        a = self.area * 1000.0
        internal = ((dd + q) * a, -dd * a) if q > 0 else (dd * a, (-dd + q) * a)
        dd_in = D / (0.5 * self.dx)            # TDispersionBC, Dirichlet face
        inflow = ((dd_in + q) * a, -dd_in * a)
        outflow = (0.0, -q * a)                # zero gradient: advection only
        return internal, inflow, outflow


def grid_faces(col: Column, T_in=None):
    """The column as pfx_grid face data (nx = ncells): internal Darcy flux
    and dispersion/distance, a Dirichlet x-min face (dispersion over half a
    cell) with boundary totals T_in, a zero-gradient x-max outflow face
    (boundary upstream, so its flux is -q).  Keyword arguments of
    runtime.DeviceGrid and the oracle's HostGrid."""
    n = col.ncells
    q = col.darcy
    D = max(col.dispersivity * abs(q) + col.saturation * col.porosity * col.tortuosity * col.diffusion, 1e-40)
    e = [np.zeros((1, 0, n)), np.zeros((0, 1, n))]
    return dict(nx=n, ny=1, nz=1, q=[np.full((1, 1, n - 1), q)] + e, dd=[np.full((1, 1, n - 1), D / col.dx)] + e,
                area=(col.area, 1.0, 1.0), bc_q=(np.full((1, 1), q), np.full((1, 1), -q)),
                bc_dd=(np.full((1, 1), D / (0.5 * col.dx)), np.zeros((1, 1))),
                bc_total=(None if T_in is None else np.asarray(T_in, dtype=np.float64), None))


def flux_residual(col: Column, T, T_in):
    """Flux part of the residual (mol/s) per cell and component; T [n][N]
    totals (mol/L), T_in [N] inflow totals.  Works on numpy and torch."""
    (tu, td), (bu, bd), (ou, od) = col.coefficients()
    r = T * 0.0
    res = tu * T[:-1] + td * T[1:]          # face i: up = i, dn = i + 1
    r[:-1] = r[:-1] + res
    r[1:] = r[1:] - res
    r[0] = r[0] - (bu * T_in + bd * T[0])
    r[-1] = r[-1] - (ou * T[-1] + od * T[-1])
    return r


def transport_step(col: Column, T, T_in, dt):
    """Explicit Euler step of the totals over dt (stable for the Courant and
    diffusion numbers of the column configuration)."""
    return T - dt * flux_residual(col, T, T_in) / (col.porosity * col.saturation * col.volume * 1000.0)


def os_steps_device_implicit(tab, dc, col: Column, T_in, dts, opts_fn):
    """Operator-split steps with the transport half implicit on the device
    (pfx_os_transport_step), then the reaction and the kinetic-state update:
    the whole step stays on the device.  Returns (per-step pfx_stats,
    per-step KrylovStats of the transport solves)."""
    from . import abi, runtime
    f = grid_faces(col, T_in)
    grid = runtime.DeviceGrid(f["nx"], f["ny"], f["nz"], f["q"], f["dd"], area=f["area"], bc_q=f["bc_q"],
                              bc_dd=f["bc_dd"], bc_total=f["bc_total"])
    tr = runtime.OsTransport(grid)
    N = int(dc.t["total"].shape[1])
    stats, tstats = [], []
    try:
        for dt in dts:
            o = abi.default_opts(dt=dt, porosity=col.porosity, saturation=col.saturation, volume=col.volume)
            rc, ts = tr.step(dc, grid, N, o)
            if rc != 0:
                raise runtime.PfxError(f"transport solve did not converge: reason {ts.reason}")
            tstats.append(ts)
            rc, st = runtime.react_batch(tab, dc, opts_fn(dt))
            if rc < 0:
                raise runtime.PfxError(f"pfx_react_batch failed: {rc}")
            runtime.update_kinetic_state(tab, dc, opts_fn(dt))
            stats.append(st)
    finally:
        tr.close()
    return stats, tstats


def os_steps_device(tab, dc, col: Column, T_in, dts, opts_fn):
    """Operator-split steps on the device: transport the totals, react,
    update the mineral volume fractions.  opts_fn(dt) gives the react
    options.  Returns the per-step pfx_stats."""
    import torch

    from . import runtime
    T_in_t = torch.as_tensor(np.asarray(T_in, dtype=np.float64), device="cuda")
    stats = []
    for dt in dts:
        dc.t["total"].copy_(transport_step(col, dc.t["total"], T_in_t, dt))
        o = opts_fn(dt)
        rc, st = runtime.react_batch(tab, dc, o)
        if rc < 0:
            raise runtime.PfxError(f"pfx_react_batch failed: {rc}")
        runtime.update_kinetic_state(tab, dc, opts_fn(dt))
        stats.append(st)
    return stats
