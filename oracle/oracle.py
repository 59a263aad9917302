"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / CPU baseline, never as
the product.  Built by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

from pflotran_amd import abi
from pflotran_amd.database import ReactionTables

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CON_TOTAL, CON_FREE, CON_LOG, CON_PH, CON_CHARGE, CON_TOTAL_SORB = 0, 1, 2, 3, 4, 5

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        lib = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp = C.c_void_p
        lib.or_react_batch.argtypes = [P(abi.ReactionDesc), P(abi.Cells), P(abi.ReactOpts), P(abi.Stats), C.c_int]
        lib.or_residual_jacobian.argtypes = [P(abi.ReactionDesc), P(abi.Cells), P(abi.ReactOpts), vp, vp, vp, C.c_int]
        lib.or_equilibrate_batch.argtypes = [P(abi.ReactionDesc), vp, vp, vp, P(abi.Cells), P(abi.ReactOpts), C.c_int,
                                             vp, C.c_int]
        lib.or_auxvar_batch.argtypes = [P(abi.ReactionDesc), P(abi.Cells), P(abi.ReactOpts), C.c_int, C.c_int]
        lib.or_update_kinetic_state.argtypes = [P(abi.ReactionDesc), P(abi.Cells), P(abi.ReactOpts), vp, C.c_int]
        lib.or_rtotal_cell.argtypes = [P(abi.ReactionDesc), vp, vp, vp, vp, C.c_double, vp, vp, vp]
        lib.or_activity_cell.argtypes = [P(abi.ReactionDesc), vp, vp, vp, vp, vp]
        lib.or_ionx_cell.argtypes = [P(abi.ReactionDesc), vp, vp, vp, vp, vp, vp]
        lib.or_ludcmp_status.argtypes = [vp, C.c_int, vp, vp]
        lib.or_ludcmp_status.restype = C.c_int
        lib.or_lubksb.argtypes = [vp, C.c_int, vp, vp]
        lib.or_logk_fit.argtypes = [vp, vp, C.c_int, vp]
        lib.or_logk_eval.argtypes = [vp, C.c_double]
        lib.or_logk_eval.restype = C.c_double
        lib.or_density_ifc67.argtypes = [C.c_double, C.c_double]
        lib.or_density_ifc67.restype = C.c_double
        lib.or_debye_params.argtypes = [C.c_double, C.c_int, vp, vp, vp]
        lib.or_gi_structure.argtypes = [P(abi.Grid), vp, vp, vp]
        lib.or_gi_dispersion.argtypes = [P(abi.Grid), vp, vp, vp, C.c_double, C.c_double, vp, vp, vp, vp]
        lib.or_gi_residual_jacobian.argtypes = [P(abi.ReactionDesc), P(abi.Cells), P(abi.Grid), P(abi.ReactOpts),
                                                vp, vp, vp, vp, C.c_int]
        _lib = lib
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


@dataclass
class HostCells:
    """numpy arrays for every pfx_cells field (cell-major)."""
    arrays: Dict[str, Optional[np.ndarray]]
    ncells: int

    @classmethod
    def allocate(cls, t: ReactionTables, n: int, sorb=None, temperature=False, per_cell_scalars=False):
        N, M, K, R = t.N, t.M, t.K, t.R
        z = lambda *s: np.zeros(s, dtype=np.float64)  # noqa: E731
        a = {
            "total": z(n, N), "total_sorb": z(n, N) if (t.has_sorption if sorb is None else sorb) else None,
            "pri_molal": np.full((n, N), 1e-9), "pri_act_coef": np.ones((n, N)),
            "sec_act_coef": np.ones((n, max(M, 1))), "sec_molal": z(n, max(M, 1)),
            "ionx_ref_sorbed": np.full((n, max(R, 1)), 1e-9), "mnrl_volfrac": z(n, max(K, 1)),
            "mnrl_area": z(n, max(K, 1)), "mnrl_rate": z(n, max(K, 1)),
            "temperature": z(n) + 25.0 if temperature else None,
            "porosity": z(n) if per_cell_scalars else None, "saturation": z(n) if per_cell_scalars else None,
            "volume": z(n) if per_cell_scalars else None, "den_kg": z(n) if per_cell_scalars else None,
            "guess": z(n, N), "num_iterations": np.zeros(n, dtype=np.int32),
            "status": np.zeros(n, dtype=np.int32), "rel_res": z(n),
        }
        return cls(a, n)

    def struct(self) -> abi.Cells:
        c = abi.Cells()
        c.ncells = self.ncells
        for f in abi.CELL_FIELDS:
            setattr(c, f, _ptr(self.arrays.get(f)))
        return c

    def copy(self) -> "HostCells":
        return HostCells({k: (None if v is None else v.copy()) for k, v in self.arrays.items()}, self.ncells)

    def __getitem__(self, k):
        return self.arrays[k]


def react_batch(t: ReactionTables, cells: HostCells, opts: abi.ReactOpts, nthreads: int = 0):
    lib = load()
    d = t.desc()
    cs = cells.struct()
    st = abi.Stats()
    rc = lib.or_react_batch(C.byref(d), C.byref(cs), C.byref(opts), C.byref(st), nthreads)
    return rc, st


def residual_jacobian(t: ReactionTables, cells: HostCells, opts: abi.ReactOpts, fixed_accum=None, nthreads=0):
    lib = load()
    d = t.desc()
    cs = cells.struct()
    n, N = cells.ncells, t.N
    res = np.zeros((n, N))
    jac = np.zeros((n, N, N))  # [cell][col][row] memory = column-major block
    rc = lib.or_residual_jacobian(C.byref(d), C.byref(cs), C.byref(opts), _ptr(fixed_accum), _ptr(res),
                                  _ptr(jac), nthreads)
    assert rc == 0
    return res, np.ascontiguousarray(jac.transpose(0, 2, 1))  # -> [cell][row][col]


def equilibrate(t: ReactionTables, ctype: Sequence[int], conc: np.ndarray, cells: HostCells,
                opts: abi.ReactOpts, molal=False, nthreads=0, rxn=None):
    """rxn: ctypes array [N] of abi.ConstraintRxn for mineral / gas rows."""
    lib = load()
    d = t.desc()
    ct = np.ascontiguousarray(ctype, dtype=np.int32)
    cc = np.ascontiguousarray(conc, dtype=np.float64).reshape(cells.ncells, t.N)
    its = np.zeros(cells.ncells, dtype=np.int32)
    cs = cells.struct()
    rc = lib.or_equilibrate_batch(C.byref(d), _ptr(ct), _ptr(cc), None if rxn is None else C.byref(rxn),
                                  C.byref(cs), C.byref(opts), int(molal), _ptr(its), nthreads)
    return rc, its


def auxvar(t: ReactionTables, cells: HostCells, opts: abi.ReactOpts, update_activity=False, nthreads=0):
    lib = load()
    d = t.desc()
    cs = cells.struct()
    lib.or_auxvar_batch(C.byref(d), C.byref(cs), C.byref(opts), int(update_activity), nthreads)


def update_kinetic_state(t: ReactionTables, cells: HostCells, opts: abi.ReactOpts, nthreads=0):
    lib = load()
    d = t.desc()
    cs = cells.struct()
    vf = np.ascontiguousarray(cells["mnrl_volfrac"])
    lib.or_update_kinetic_state(C.byref(d), C.byref(cs), C.byref(opts), _ptr(vf), nthreads)
    cells.arrays["mnrl_volfrac"] = vf


def batch_steps(t: ReactionTables, cells: HostCells, opts: abi.ReactOpts, dts: Sequence[float], nthreads=1):
    """Global-implicit batch run without transport: per step the reference
    refreshes the fixed accumulation from the current state
    (RTUpdateFixedAccumulation, reactive_transport.F90:726-848), then updates
    activity coefficients (PMRTInitializeTimestep, pm_rt.F90:410-413), solves
    the per-cell residual, and updates mineral volume fractions
    (RUpdateKineticState, reaction.F90:5340-5366).  The per-cell solve is the
    same system as the operator-split RReact with dt-consistent units."""
    its = []
    for dt in dts:
        opts.dt = dt
        auxvar(t, cells, opts, update_activity=False, nthreads=nthreads)   # totals at the old state
        rc, st = react_batch(t, cells, opts, nthreads=nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle react failed rc={rc} stats={st.as_dict()}")
        its.append(st.total_iterations)
        update_kinetic_state(t, cells, opts, nthreads=nthreads)
    auxvar(t, cells, opts, update_activity=False, nthreads=nthreads)
    return its


def lu_solve(A: np.ndarray, b: np.ndarray):
    """ludcmp/lubksb on a copy (column-major as the reference)."""
    lib = load()
    n = A.shape[0]
    a = np.asfortranarray(A.astype(np.float64)).copy(order="F")
    indx = np.zeros(n, dtype=np.int32)
    dd = np.zeros(1, dtype=np.int32)
    st = lib.or_ludcmp_status(a.ctypes.data, n, indx.ctypes.data, dd.ctypes.data)
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    if st == 0:
        lib.or_lubksb(a.ctypes.data, n, indx.ctypes.data, x.ctypes.data)
    return st, x, indx, a


def density_ifc67(t=25.0, p=101325.0) -> float:
    return load().or_density_ifc67(t, p)


def debye(tref=25.0, use_bdot=True):
    A, B, D = (np.zeros(1) for _ in range(3))
    load().or_debye_params(tref, int(use_bdot), A.ctypes.data, B.ctypes.data, D.ctypes.data)
    return float(A[0]), float(B[0]), float(D[0])


def logk_fit(temps, logK):
    t = np.ascontiguousarray(temps, dtype=np.float64)
    k = np.ascontiguousarray(logK, dtype=np.float64)
    c = np.zeros(5)
    load().or_logk_fit(t.ctypes.data, k.ctypes.data, len(t), c.ctypes.data)
    return c


def speciated_state(t, n, seed=None, workload_id=2, temperature=True, nthreads=0, only_ok=True):
    """Constraint speciation (totals of every primary but H+, pH) of seeded
    synthetic cells by the oracle: the previous state of each cell before
    its reaction step.  With only_ok, cells whose speciation fails are
    replaced by further draws."""
    from pflotran_amd import synth
    if seed is None:
        seed = synth.SEED
    if only_ok:
        cells, ctype, conc, ok = speciated_state(t, int(n * 1.2) + 8, seed, workload_id, temperature, nthreads,
                                                 only_ok=False)
        idx = np.where(ok)[0][:n]
        assert len(idx) == n, "not enough speciated cells"
        sub = HostCells({k: (None if v is None else v[idx].copy()) for k, v in cells.arrays.items()}, n)
        return sub, ctype, conc[idx], np.ones(n, dtype=bool)
    ci = synth.random_cells(t, n, seed=seed, workload_id=workload_id)
    cells = HostCells.allocate(t, n, temperature=temperature)
    if temperature:
        cells["temperature"][:] = ci.temp
    cells["mnrl_volfrac"][:] = ci.vf[:, :cells["mnrl_volfrac"].shape[1]]
    cells["mnrl_area"][:] = ci.area[:, :cells["mnrl_area"].shape[1]]
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    ctype = [CON_PH] + [CON_TOTAL] * (t.N - 1)
    rc, its = equilibrate(t, ctype, conc, cells, synth.react_opts(), nthreads=nthreads)
    ok = (its > 0) & np.all(np.isfinite(cells["pri_molal"]), axis=1) & np.all(cells["pri_molal"] > 0, axis=1)
    return cells, ctype, conc, ok


class HostGrid:
    """numpy arrays of a pfx_grid (structured box, natural ordering)."""

    def __init__(self, nx, ny, nz, q, dd, area=(1.0, 1.0, 1.0), bc_q=(None, None), bc_dd=(None, None),
                 bc_total=(None, None), log_formulation=False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.q = [np.ascontiguousarray(a, dtype=np.float64) for a in q]
        self.dd = [np.ascontiguousarray(a, dtype=np.float64) for a in dd]
        self.area = tuple(float(a) for a in area)
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        self.bc_q = [conv(a) for a in bc_q]
        self.bc_dd = [conv(a) for a in bc_dd]
        self.bc_total = [conv(a) for a in bc_total]
        self.log_formulation = bool(log_formulation)

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    def struct(self) -> abi.Grid:
        g = abi.Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.log_formulation = int(self.log_formulation)
        for d in range(3):
            g.area[d] = self.area[d]
            g.q[d] = _ptr(self.q[d])
            g.dd[d] = _ptr(self.dd[d])
        for s_ in range(2):
            g.bc_q[s_] = _ptr(self.bc_q[s_])
            g.bc_dd[s_] = _ptr(self.bc_dd[s_])
            g.bc_total[s_] = _ptr(self.bc_total[s_])
        return g


def gi_structure(grid: HostGrid):
    lib = load()
    g = grid.struct()
    nnzb = C.c_int64(0)
    lib.or_gi_structure(C.byref(g), None, None, C.byref(nnzb))
    row_ptr = np.zeros(grid.ncells + 1, dtype=np.int32)
    col_idx = np.zeros(nnzb.value, dtype=np.int32)
    lib.or_gi_structure(C.byref(g), row_ptr.ctypes.data, col_idx.ctypes.data, C.byref(nnzb))
    return row_ptr, col_idx


def gi_residual_jacobian(t: ReactionTables, cells: HostCells, grid: HostGrid, opts: abi.ReactOpts, nthreads=0):
    """Global residual [n, N] and BSR values [nnzb, N, N] (row, col) with
    (row_ptr, col_idx)."""
    lib = load()
    row_ptr, col_idx = gi_structure(grid)
    n, N = grid.ncells, t.N
    r = np.zeros((n, N))
    vals = np.zeros((len(col_idx), N, N))   # [block][col][row] memory = column-major blocks
    d = t.desc()
    cs = cells.struct()
    g = grid.struct()
    rc = lib.or_gi_residual_jacobian(C.byref(d), C.byref(cs), C.byref(g), C.byref(opts), row_ptr.ctypes.data,
                                     col_idx.ctypes.data, r.ctypes.data, vals.ctypes.data, nthreads)
    if rc != 0:
        raise RuntimeError(f"or_gi_residual_jacobian failed: {rc}")
    return r, np.ascontiguousarray(vals.transpose(0, 2, 1)), row_ptr, col_idx


def bsr_to_dense(vals, row_ptr, col_idx):
    """Dense [n*N, n*N] matrix of BSR blocks given as [block][row][col]."""
    nb, N, _ = vals.shape
    n = len(row_ptr) - 1
    A = np.zeros((n * N, n * N))
    for row in range(n):
        for b in range(row_ptr[row], row_ptr[row + 1]):
            col = col_idx[b]
            A[row * N:(row + 1) * N, col * N:(col + 1) * N] = vals[b]
    return A


def gi_dispersion(grid: HostGrid, porosity, saturation=None, tortuosity=None, dispersivity=0.0, diffusion=1e-9,
                  dx=(1.0, 1.0, 1.0), bc_type=(0, 0)):
    """Face dispersion coefficients (internal [3] arrays, boundary [2])."""
    lib = load()
    g = grid.struct()
    dd = [np.zeros_like(a) for a in grid.q]
    bc = [np.zeros((grid.nz, grid.ny)) for _ in range(2)]
    ddp = (C.c_void_p * 3)(*[a.ctypes.data for a in dd])
    bcp = (C.c_void_p * 2)(*[a.ctypes.data for a in bc])
    conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    phi, sat, tau = conv(porosity), conv(saturation), conv(tortuosity)
    dxa = np.ascontiguousarray(dx, dtype=np.float64)
    bt = np.ascontiguousarray(bc_type, dtype=np.int32)
    lib.or_gi_dispersion(C.byref(g), _ptr(phi), _ptr(sat), _ptr(tau), dispersivity, diffusion, dxa.ctypes.data,
                         bt.ctypes.data, ddp, bcp)
    return dd, bc



# --------------------------------------------------------------------------
# operator-split transport half (numpy/scipy restatement)


def _flux_coef(area, q, dd):
    """TFluxCoef with upwinding (transport.F90:756-819): (T_up, T_dn) in L/s."""
    up = np.where(q > 0.0, dd + q, dd)
    dn = np.where(q > 0.0, -dd, -dd + q)
    return up * area * 1000.0, dn * area * 1000.0


def os_transport_matrix(grid: HostGrid, acc: np.ndarray):
    """Scalar transport matrix of RTCalculateTransportMatrix
    (reactive_transport.F90:1379-1601) as a scipy CSR matrix, natural cell
    order: per internal face (up, dn) A[up,up] += T_up, A[up,dn] += T_dn,
    A[dn,dn] -= T_dn, A[dn,up] -= T_up (:1442-1486); per boundary face
    A[c,c] -= T_dn (:1489-1518), and -= T_up as well when the face carries
    the cell's own total (zero gradient); accumulation phi*s*1000*V/dt on the
    diagonal (:1526-1532)."""
    import scipy.sparse as sp
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    n = nx * ny * nz
    cid = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [np.asarray(acc, dtype=np.float64).ravel()]
    slices = [(np.s_[:, :, :-1], np.s_[:, :, 1:]), (np.s_[:, :-1, :], np.s_[:, 1:, :]),
              (np.s_[:-1, :, :], np.s_[1:, :, :])]
    for d in range(3):
        if grid.q[d].size == 0:
            continue
        tu, td = _flux_coef(grid.area[d], grid.q[d], grid.dd[d])
        up, dn = cid[slices[d][0]].ravel(), cid[slices[d][1]].ravel()
        tu, td = tu.ravel(), td.ravel()
        rows += [up, up, dn, dn]
        cols += [up, dn, dn, up]
        vals += [tu, td, -td, -tu]
    for side in range(2):
        if grid.bc_q[side] is None:
            continue
        tu, td = _flux_coef(grid.area[0], grid.bc_q[side], grid.bc_dd[side])
        c = cid[:, :, 0 if side == 0 else nx - 1].ravel()
        diag = -td.ravel() - (tu.ravel() if grid.bc_total[side] is None else 0.0)
        rows.append(c)
        cols.append(c)
        vals.append(diag)
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A.tocsr()


def os_transport_rhs(grid: HostGrid, acc: np.ndarray, total: np.ndarray):
    """RTCalculateRHS_t0 (reactive_transport.F90:1085-1146): acc * T_old,
    plus RTCalculateRHS_t1's T_up * T_bc on boundary faces with a boundary
    total (:1232-1258); [ncells][N]."""
    b = np.asarray(acc, dtype=np.float64).reshape(-1, 1) * total
    nx = grid.nx
    cid = np.arange(grid.ncells).reshape(grid.nz, grid.ny, nx)
    for side in range(2):
        if grid.bc_q[side] is None or grid.bc_total[side] is None:
            continue
        tu, _ = _flux_coef(grid.area[0], grid.bc_q[side], grid.bc_dd[side])
        c = cid[:, :, 0 if side == 0 else nx - 1].ravel()
        b[c] += tu.ravel()[:, None] * grid.bc_total[side][None, :]
    return b


def os_transport_step(grid: HostGrid, acc: np.ndarray, total: np.ndarray):
    """One implicit transport step of every component (direct solve)."""
    import scipy.sparse.linalg as spla
    A = os_transport_matrix(grid, acc).tocsc()
    b = os_transport_rhs(grid, acc, total)
    lu = spla.splu(A)
    return np.stack([lu.solve(b[:, j]) for j in range(b.shape[1])], axis=1)


# --------------------------------------------------------------------------
# global-implicit column run (numpy/scipy restatement of the time loop)


def gi_newton(t: ReactionTables, cells: HostCells, grid: HostGrid, opts: abi.ReactOpts, rtol=1e-8, stol=1e-8,
              atol=1e-50, max_it=50, max_dlnC=5.0, itol_res=None, itol_upd=None, rel_update=None,
              max_norm=1e20):
    """PETSc SNES NEWTONLS with the basic line search on the oracle's
    residual and Jacobian, each linear system solved directly (the
    reference's ILU(0) on a 1-D block-tridiagonal matrix, or its DIRECT
    solver, is exact): PMRTCheckUpdatePre's clip of the log-form update
    (pm_rt.F90:734-840), then the tests of one Newton iteration:
    SNESConvergedDefault in its order (NaN, ||F|| < atol, and from the first
    update on ||F|| <= rtol ||F0||, ||y|| < stol ||x||), PFLOTRAN's
    PMRTCheckUpdatePost (pm_rt.F90:885-925: with ITOL_RELATIVE_UPDATE,
    max |y / x_old| below it converges, reason 12) and ConvergenceTest
    (convergence.F90:170-245: ||F||_inf below ITOL converges, reason 10;
    ||y||_inf below ITOL_UPDATE, 11; ||F||_inf above max_norm diverges, -10;
    no decision at iteration 0, newton_min_iterations 1), then the
    iteration limit.  Updates cells["pri_molal"]; returns (iterations,
    reason)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    N = t.N
    n = grid.ncells
    r, v, rp, ci = gi_residual_jacobian(t, cells, grid, opts)
    f0 = np.linalg.norm(r)
    if not np.isfinite(f0):
        return 0, abi.PFX_SNES_DIVERGED_FNORM_NAN
    it = 0
    while True:
        if it >= max_it:
            return it, abi.PFX_SNES_DIVERGED_MAX_IT
        A = sp.bsr_matrix((v, ci, rp), shape=(n * N, n * N)).tocsc()
        y = spla.spsolve(A, r.ravel())
        x_old = np.log(cells["pri_molal"].ravel()) if grid.log_formulation else cells["pri_molal"].ravel().copy()
        if grid.log_formulation:
            y = np.sign(y) * np.minimum(np.abs(y), max_dlnC)
            m = cells["pri_molal"].ravel() * np.exp(-y)
        else:
            m0 = cells["pri_molal"].ravel()
            with np.errstate(over="ignore"):
                ratio = np.where(m0 <= y, np.abs(m0 / np.where(y == 0, 1.0, y)), 1e20)
            mr = ratio.min() if ratio.size else 1e20
            if mr < 1.0:
                y = y * mr * 0.99
            m = m0 - y
        cells["pri_molal"][:] = m.reshape(n, N)
        it += 1
        r, v, rp, ci = gi_residual_jacobian(t, cells, grid, opts)
        f = np.linalg.norm(r)
        x = np.log(m) if grid.log_formulation else m
        reason = 0
        if not np.isfinite(f):
            reason = abi.PFX_SNES_DIVERGED_FNORM_NAN
        elif f < atol:
            reason = abi.PFX_SNES_CONVERGED_FNORM_ABS
        elif f <= rtol * f0:
            reason = abi.PFX_SNES_CONVERGED_FNORM_RELATIVE
        elif np.linalg.norm(y) < stol * np.linalg.norm(x):
            reason = abi.PFX_SNES_CONVERGED_SNORM_RELATIVE
        if rel_update is not None:
            with np.errstate(divide="ignore", invalid="ignore"):
                if np.max(np.abs(y / x_old)) < rel_update:
                    reason = abi.PFX_SNES_CONVERGED_RELATIVE_UPDATE
        fi = np.max(np.abs(r)) if r.size else 0.0
        if itol_res is not None and fi < itol_res:
            reason = abi.PFX_SNES_CONVERGED_ITOL_RES
        if itol_upd is not None and np.max(np.abs(y)) < itol_upd:
            reason = abi.PFX_SNES_CONVERGED_ITOL_UPDATE
        if fi > max_norm:
            reason = abi.PFX_SNES_DIVERGED_MAX_NORM
        if reason != 0:
            return it, reason


def gi_column_run(deck, db=None, max_cuts=16):
    """The deck's 1-D column by the global-implicit time loop on the CPU,
    in the order of pflotran_amd.gi_run.ColumnRun (the reference's
    PMRTInitializeTimestep, SNESSolve with the time-step cut,
    PMRTUpdateSolution, TimestepperBEUpdateDT), the step controller being
    pflotran_amd.stepper.  Returns (stepper, cells, tables)."""
    from pflotran_amd.gi_run import face_data
    db = db if db is not None else deck.database()
    t = deck.tables(db)
    box = deck.box()
    den = deck.water_density()
    n = box.nx

    def opts(dt):
        return abi.default_opts(dt=dt, porosity=box.porosity, saturation=1.0, volume=box.volume, den_kg=den)

    def speciate(name, ncell):
        setup = deck.constraint_setup(name, t, db)
        c = HostCells.allocate(t, 1)
        if t.K and setup.get("vf") is not None:
            c["mnrl_volfrac"][0, :] = setup["vf"]
            c["mnrl_area"][0, :] = setup["area"]
        o = opts(1.0)
        if setup.get("guess") is not None:
            c["guess"][0, :] = setup["guess"]
            o.guess_mode = abi.PFX_GUESS_ARRAY
        rc, its = equilibrate(t, setup["ctype"], np.asarray([setup["conc"]], dtype=np.float64), c, o,
                              molal=bool(setup.get("molal", 0)), rxn=setup.get("rxn"))
        if rc != 0:
            raise RuntimeError(f"constraint {name} did not speciate")
        c = HostCells({k: (None if v is None else np.repeat(v, ncell, axis=0)) for k, v in c.arrays.items()}, ncell)
        c["pri_act_coef"][:] = 1.0
        if c.arrays.get("sec_act_coef") is not None:
            c["sec_act_coef"][:] = 1.0
        if c.arrays.get("ionx_ref_sorbed") is not None:
            c["ionx_ref_sorbed"][:] = 1e-9
        auxvar(t, c, opts(1.0), update_activity=False)
        return c

    cells = speciate(box.initial, n)
    if t.act_freq != abi.PFX_ACT_OFF:
        auxvar(t, cells, opts(1.0), update_activity=True)
        auxvar(t, cells, opts(1.0), update_activity=True)
    qf, ddf, bc_q, bc_dd, dirichlet = face_data(box)
    bcs = [speciate(getattr(box, f)[1], 1) if dirichlet[i] else None for i, f in enumerate(("west", "east"))]
    bc_total = [None if c is None else c["total"][0].copy() for c in bcs]
    grid = HostGrid(n, 1, 1, qf, ddf, area=(box.area, 1.0, 1.0), bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total,
                    log_formulation=bool(t.use_log_formulation))
    ts = deck.stepper()
    act = t.act_freq == abi.PFX_ACT_TIMESTEP
    while not ts.done:
        dt = ts.set_target_time()
        if act:
            c2 = cells.copy()
            auxvar(t, c2, opts(1.0), update_activity=True)
            for k in ("pri_act_coef", "sec_act_coef", "sec_molal"):
                if cells.arrays.get(k) is not None:
                    cells[k][:] = c2[k]
            for i, c in enumerate(bcs):
                if c is not None:
                    auxvar(t, c, opts(1.0), update_activity=True)
                    grid.bc_total[i][:] = c["total"][0]
        old = {k: cells[k].copy() for k in ("pri_molal", "ionx_ref_sorbed") if cells.arrays.get(k) is not None}
        cuts, wasted, dt_try = 0, 0, dt
        nw = deck.newton_options()
        while True:
            its, reason = gi_newton(t, cells, grid, opts(dt_try), **nw)
            if reason > 0:
                break
            wasted += its
            if cuts >= max_cuts:
                raise RuntimeError(f"step {ts.steps + 1} failed after {cuts} cuts")
            cuts += 1
            dt_try *= 0.5
            for k, v in old.items():
                cells[k][:] = v
        ts.accept(dt_try, its, cuts)
        o = opts(dt_try)
        if t.K:
            update_kinetic_state(t, cells, o)
        auxvar(t, cells, o, update_activity=False)
    return ts, cells, t
