"""Device-resident batches of synthetic cells at the start of a reaction step.

A cell's previous state is its constraint speciation (pH for H+, total
concentration for every other primary, pfx_equilibrate_batch) followed by the
initial aux-var refresh of the reference (unit activity coefficients, then two
activity passes: condition_control.F90:929-935).  Cells whose speciation fails
are replaced by further draws, so a batch holds exactly the cells asked for.
"""
from __future__ import annotations

from . import abi, runtime, synth
from .database import ReactionTables

# fields a reaction step overwrites: restore them to repeat the same step
STEP_INPUTS = ("total", "total_sorb", "pri_molal", "pri_act_coef", "sec_act_coef", "ionx_ref_sorbed")


def speciated_cells(t: ReactionTables, tab: runtime.Tables, n: int, seed: int = synth.SEED,
                    workload_id: int = 2, oversample: float = 1.08):
    """n speciated cells on the device (runtime.DeviceCells) and the number
    of failed draws."""
    import torch
    n_draw = int(n * oversample) + 64
    ci = synth.random_cells(t, n_draw, seed=seed, workload_id=workload_id)
    arr = synth.host_state_arrays(t, n_draw, with_temperature=True)
    arr["temperature"][:] = ci.temp
    arr["mnrl_volfrac"][:] = ci.vf
    arr["mnrl_area"][:] = ci.area
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    ctype = [abi.PFX_CONSTRAINT_PH] + [abi.PFX_CONSTRAINT_TOTAL] * (t.N - 1)
    dc = runtime.DeviceCells.from_host(arr, n_draw)
    runtime.equilibrate(tab, dc, ctype, conc, synth.react_opts(max_iterations=300))
    m = dc.t["pri_molal"]
    ok = (dc.t["status"] == 0) & torch.isfinite(m).all(dim=1) & (m > 0).all(dim=1)
    idx = torch.nonzero(ok).flatten()[:n]
    if idx.numel() < n:
        raise runtime.PfxError(f"only {idx.numel()} of {n} cells speciated")
    dc = runtime.DeviceCells({k: (None if v is None else v.index_select(0, idx).contiguous())
                              for k, v in dc.t.items()}, n)
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    dc.t["ionx_ref_sorbed"].fill_(1e-9)
    runtime.auxvar(tab, dc, 2, synth.react_opts())
    return dc, n_draw - int(ok.sum())


def snapshot(dc: runtime.DeviceCells):
    return {k: dc.t[k].clone() for k in STEP_INPUTS if dc.t.get(k) is not None}


def restore(dc: runtime.DeviceCells, snap):
    for k, v in snap.items():
        dc.t[k].copy_(v)


def host_arrays(dc: runtime.DeviceCells):
    return {k: (None if v is None else v.cpu().numpy()) for k, v in dc.t.items()}


def subset(dc: runtime.DeviceCells, idx) -> runtime.DeviceCells:
    return runtime.DeviceCells({k: (None if v is None else v.index_select(0, idx).contiguous())
                                for k, v in dc.t.items()}, int(idx.numel()))


__all__ = ["speciated_cells", "snapshot", "restore", "host_arrays", "subset", "STEP_INPUTS"]


# ---------------------------------------------------------------------------
# structured grid of the global-implicit configuration
D_MOLECULAR = 1e-9      # m^2/s
TORTUOSITY = 1.0
POROSITY = 0.25
Q_MEAN = 3.1688e-8      # 1 m/yr


def gi_grid(nx: int, ny: int, nz: int, seed: int = synth.SEED, workload_id: int = 5, dx: float = 1.0,
            log_formulation: bool = True, bc_total=None):
    """Host arrays of a 7-point grid with a seeded log-normal permeability
    (ln k ~ N(ln 1e-12, 1)) and a steady flow field: each x row carries the
    1-D Darcy flux of its harmonic-mean permeability under one head drop
    for the whole box, scaled to a mean of 1 m/yr (divergence-free; no
    cross flow).  Dispersion: saturation * porosity * tortuosity * D_m over
    the cell distance (alpha_L = 0, TDispersion); at the Dirichlet x-min
    inflow over dx/2, none at the zero-gradient x-max outflow
    (TDispersionBC).  Returns a dict of numpy arrays for runtime.DeviceGrid."""
    import numpy as np
    lnk = synth.rng(workload_id, "perm").normal(np.log(1e-12), 1.0, size=(nz, ny, nx))
    k = np.exp(lnk)
    k_h = nx / np.sum(1.0 / k, axis=2)                 # [nz][ny]
    q_row = Q_MEAN * k_h / np.exp(np.mean(np.log(k_h)))
    dd = POROSITY * 1.0 * TORTUOSITY * D_MOLECULAR / dx
    q = [np.repeat(q_row[:, :, None], max(nx - 1, 0), axis=2), np.zeros((nz, max(ny - 1, 0), nx)),
         np.zeros((max(nz - 1, 0), ny, nx))]
    dda = [np.full(a.shape, dd) for a in q]
    return dict(nx=nx, ny=ny, nz=nz, q=q, dd=dda, area=(dx * dx,) * 3,
                bc_q=(q_row.copy(), -q_row.copy()), bc_dd=(np.full((nz, ny), 2 * dd), np.zeros((nz, ny))),
                bc_total=(bc_total, None), log_formulation=log_formulation)


# ---------------------------------------------------------------------------
# uneven batch: chemistry zones, half the cells near a mineral's saturation
ZONE_N = (8, 12, 20, 30)
ZONE_CELLS = (1_000_000, 500_000, 300_000, 200_000)


def zone_networks(seed: int = synth.SEED):
    """One network per zone: N primaries, 3N complexes, 3 kinetic minerals."""
    return [synth.random_network(N, 3 * N, 3, ionx=False, seed=seed + 101 * N) for N in ZONE_N]


def mineral_log_omega(t: ReactionTables, dc: runtime.DeviceCells, k: int = 0):
    """log10 saturation index of mineral k at the cells' state (host)."""
    import numpy as np

    from .database import logk_eval
    m = dc.t["pri_molal"].cpu().numpy()
    g = dc.t["pri_act_coef"].cpu().numpy()
    temp = dc.t["temperature"].cpu().numpy() if dc.t.get("temperature") is not None else np.full(len(m), 25.0)
    la = np.log10(m * g)
    ns = int(t.mnrl_specid[k, 0])
    lq = np.zeros(len(m))
    for s_ in range(ns):
        lq += t.mnrl_stoich[k, s_] * la[:, t.mnrl_specid[k, 1 + s_]]
    if t.mnrl_logKcoef is not None:
        lk = logk_eval(t.mnrl_logKcoef[k][None, :], temp)
    else:
        lk = np.full(len(m), t.mnrl_logK[k])
    return lq - lk


def near_saturation_cells(t: ReactionTables, tab: runtime.Tables, n: int, seed: int, workload_id: int,
                          fraction: float = 0.5, window: float = 0.05, passes: int = 3, oversample: float = 2.2):
    """n speciated cells of which exactly round(fraction * n) lie within
    |log Omega| <= window of mineral 0 and the rest are ordinary draws.
    The first 60% of the oversampled draws are moved toward a target in the
    window by rescaling the total of mineral 0's dominant non-H+ primary and
    re-speciating (a few passes); after the initial aux-var refresh (two
    activity passes, as the reaction step sees the cells) the cells are
    taken in draw order: the first in-window cells of the moved draws and the
    first cells of the others.  Returns the cells, their log Omega and the
    mask of cells inside the window."""
    import numpy as np
    import torch
    n_draw = int(n * oversample) + 64
    ci = synth.random_cells(t, n_draw, seed=seed, workload_id=workload_id)
    arr = synth.host_state_arrays(t, n_draw, with_temperature=True)
    arr["temperature"][:] = ci.temp
    arr["mnrl_volfrac"][:] = ci.vf
    arr["mnrl_area"][:] = ci.area
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    ctype = [abi.PFX_CONSTRAINT_PH] + [abi.PFX_CONSTRAINT_TOTAL] * (t.N - 1)
    # dominant non-H+ primary of mineral 0 and its coefficient
    ns = int(t.mnrl_specid[0, 0])
    cand = [(abs(t.mnrl_stoich[0, s_]), int(t.mnrl_specid[0, 1 + s_]), t.mnrl_stoich[0, s_]) for s_ in range(ns)
            if int(t.mnrl_specid[0, 1 + s_]) != 0]
    _, jp, nu = max(cand)
    moved = np.zeros(n_draw, dtype=bool)
    moved[:int(0.6 * n_draw)] = True
    target = synth.rng(workload_id, "perturb", seed).uniform(-window, window, size=n_draw)
    dc = None
    for it in range(passes + 1):
        dc = runtime.DeviceCells.from_host(arr, n_draw)
        runtime.equilibrate(tab, dc, ctype, conc, synth.react_opts(max_iterations=300))
        if it == passes:
            break
        lo = mineral_log_omega(t, dc)
        shift = np.where(moved & np.isfinite(lo), (target - lo) / nu, 0.0)
        conc[:, jp] *= 10.0 ** np.clip(shift, -2.0, 2.0)
        conc[:, jp] = np.clip(conc[:, jp], 1e-12, 1.0)
    m = dc.t["pri_molal"]
    ok = ((dc.t["status"] == 0) & torch.isfinite(m).all(dim=1) & (m > 0).all(dim=1)).cpu().numpy()
    # the initial refresh of every draw, then the selection
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    dc.t["ionx_ref_sorbed"].fill_(1e-9)
    runtime.auxvar(tab, dc, 2, synth.react_opts())
    lo_all = mineral_log_omega(t, dc)
    inside_all = ok & np.isfinite(lo_all) & (np.abs(lo_all) <= window)
    n_near = int(round(fraction * n))
    sel_near = np.nonzero(moved & inside_all)[0][:n_near]
    sel_far = np.nonzero(~moved & ok)[0][:n - n_near]
    if len(sel_near) < n_near or len(sel_far) < n - n_near:
        raise runtime.PfxError(f"only {len(sel_near)} near-saturation and {len(sel_far)} other cells of {n}")
    idx = np.sort(np.concatenate([sel_near, sel_far]))
    dc = subset(dc, torch.as_tensor(idx, device=dc.t["pri_molal"].device))
    lo = lo_all[idx]
    inside = inside_all[idx] & moved[idx]
    return dc, lo, inside
