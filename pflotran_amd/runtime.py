"""Python mirror of the C-ABI (libpfx_chem.so): device state and batch calls.

PyTorch is used only for device memory and streams; all chemistry runs in the
HIP kernels of libpfx_chem.so.  There is no CPU fallback: if the library is
missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

from . import abi
from .database import ReactionTables

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfx_chem.so")

EXPORTED = ["pfx_abi_version", "pfx_tables_create", "pfx_tables_destroy", "pfx_react_batch", "pfx_tables_stats",
            "pfx_react_batch_noalloc",
            "pfx_residual_jacobian", "pfx_equilibrate_batch", "pfx_auxvar_batch", "pfx_update_kinetic_state",
            "pfx_gi_structure", "pfx_gi_residual_jacobian", "pfx_tables_timing", "pfx_gi_dispersion",
            "pfx_react_zones", "pfx_zone_plan_rebalance", "pfx_zones_graph_create", "pfx_graph_launch",
            "pfx_graph_stats", "pfx_graph_destroy", "pfx_tables_tune",
            # learned warm start (include/pfx_surrogate.h)
            "pfx_mlp_create", "pfx_mlp_destroy", "pfx_mlp_params", "pfx_mlp_forward", "pfx_mlp_gradients",
            "pfx_mlp_adam", "pfx_mlp_train_step", "pfx_surrogate_features", "pfx_surrogate_guess",
            "pfx_surrogate_guess_from",
            # global-implicit step on the device (include/pfx_gi_solver.h)
            "pfx_bsr_create", "pfx_bsr_destroy", "pfx_bsr_spmv", "pfx_bsr_solve", "pfx_gi_solver_create",
            "pfx_gi_solver_destroy", "pfx_gi_solver_step", "pfx_gi_solver_timestep", "pfx_os_transport_step",
            "pfx_bsr_solve_multi", "pfx_bsr_multi_stats", "pfx_os_graph_create", "pfx_os_graph_launch",
            "pfx_os_graph_stats", "pfx_os_graph_destroy"]

_lib = None


class PfxError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libpfx_chem.so and declare the C signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PfxError(f"libpfx_chem.so not built: {path} (run `python -m pflotran_amd.build`)")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname); loading torch first makes the library bind to that copy
    # instead of pulling a second runtime from /opt/rocm.
    import torch  # noqa: F401
    lib = C.CDLL(path)
    P = C.POINTER
    lib.pfx_abi_version.restype = C.c_int
    lib.pfx_tables_create.argtypes = [P(abi.ReactionDesc), P(C.c_void_p)]
    lib.pfx_tables_destroy.argtypes = [C.c_void_p]
    lib.pfx_react_batch.argtypes = [C.c_void_p, P(abi.Cells), P(abi.ReactOpts), P(abi.Stats)]
    lib.pfx_tables_stats.argtypes = [C.c_void_p, P(abi.Stats)]
    lib.pfx_react_batch_noalloc.argtypes = [C.c_void_p, P(abi.Cells), P(abi.ReactOpts), P(abi.Stats)]
    lib.pfx_residual_jacobian.argtypes = [C.c_void_p, P(abi.Cells), P(abi.ReactOpts), C.c_void_p, C.c_void_p]
    lib.pfx_equilibrate_batch.argtypes = [C.c_void_p, P(abi.Cells), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                          P(abi.ReactOpts), P(abi.Stats)]
    lib.pfx_auxvar_batch.argtypes = [C.c_void_p, P(abi.Cells), C.c_int32, P(abi.ReactOpts)]
    lib.pfx_update_kinetic_state.argtypes = [C.c_void_p, P(abi.Cells), P(abi.ReactOpts)]
    lib.pfx_tables_timing.argtypes = [C.c_void_p, C.c_void_p]
    lib.pfx_tables_tune.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.pfx_react_zones.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, P(abi.ReactOpts), P(abi.ZonePlan),
                                    C.c_void_p]
    lib.pfx_zone_plan_rebalance.argtypes = [P(abi.ZonePlan)]
    lib.pfx_zones_graph_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, P(abi.ReactOpts), P(abi.ZonePlan),
                                           P(C.c_void_p)]
    lib.pfx_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
    lib.pfx_graph_stats.argtypes = [C.c_void_p, C.c_void_p, P(abi.ZonePlan)]
    lib.pfx_graph_destroy.argtypes = [C.c_void_p]
    lib.pfx_gi_structure.argtypes = [P(abi.Grid), C.c_void_p, C.c_void_p, P(C.c_int64)]
    lib.pfx_gi_dispersion.argtypes = [P(abi.Grid), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.pfx_gi_residual_jacobian.argtypes = [C.c_void_p, P(abi.Cells), P(abi.Grid), P(abi.ReactOpts), C.c_void_p,
                                             C.c_void_p, C.c_void_p]
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    lib.pfx_mlp_create.argtypes = [i32, vp, i32, C.c_uint64, P(vp)]
    lib.pfx_mlp_destroy.argtypes = [vp]
    lib.pfx_mlp_params.argtypes = [vp, i32, i32, vp, i32]
    lib.pfx_mlp_forward.argtypes = [vp, vp, i64, vp, vp]
    lib.pfx_mlp_gradients.argtypes = [vp, vp, vp, i32, vp, vp]
    lib.pfx_mlp_adam.argtypes = [vp, f32, f32, f32, f32, vp]
    lib.pfx_mlp_train_step.argtypes = [vp, vp, vp, i32, f32, vp, vp]
    lib.pfx_surrogate_features.argtypes = [vp, vp, P(abi.Cells), vp, vp]
    lib.pfx_surrogate_guess.argtypes = [vp, vp, P(abi.Cells), vp]
    lib.pfx_surrogate_guess_from.argtypes = [vp, vp, i64, i32, vp, vp, vp]
    lib.pfx_bsr_create.argtypes = [i64, i32, i64, vp, vp, P(vp)]
    lib.pfx_bsr_destroy.argtypes = [vp]
    lib.pfx_bsr_spmv.argtypes = [vp, vp, vp, vp, vp]
    lib.pfx_bsr_solve.argtypes = [vp, vp, vp, vp, P(abi.KrylovOpts), vp, P(abi.KrylovStats)]
    lib.pfx_gi_solver_create.argtypes = [vp, P(abi.Grid), P(vp)]
    lib.pfx_gi_solver_destroy.argtypes = [vp]
    lib.pfx_gi_solver_step.argtypes = [vp, P(abi.Cells), P(abi.Grid), P(abi.ReactOpts), P(abi.NewtonOpts),
                                       P(abi.NewtonStats)]
    lib.pfx_gi_solver_timestep.argtypes = [vp, P(abi.Cells), P(abi.Grid), P(abi.ReactOpts), P(abi.NewtonOpts), i32,
                                           C.c_double, P(abi.GiStepStats)]
    lib.pfx_os_transport_step.argtypes = [vp, P(abi.Grid), P(abi.Cells), i32, P(abi.ReactOpts), vp,
                                          P(abi.KrylovOpts), P(abi.KrylovStats)]
    lib.pfx_bsr_solve_multi.argtypes = [vp, vp, vp, vp, i32, P(abi.KrylovOpts), vp, i32, P(abi.KrylovStats)]
    lib.pfx_bsr_multi_stats.argtypes = [vp, vp, P(abi.KrylovStats)]
    lib.pfx_os_graph_create.argtypes = [vp, P(abi.Grid), vp, P(abi.Cells), i32, P(abi.ReactOpts), P(abi.ReactOpts),
                                        vp, P(abi.KrylovOpts), P(vp)]
    lib.pfx_os_graph_launch.argtypes = [vp, vp]
    lib.pfx_os_graph_stats.argtypes = [vp, P(i32), P(abi.Stats), P(abi.KrylovStats)]
    lib.pfx_os_graph_destroy.argtypes = [vp]
    if lib.pfx_abi_version() != abi.PFX_ABI_VERSION:
        raise PfxError("ABI version mismatch between libpfx_chem.so and pflotran_amd.abi")
    _lib = lib
    return lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise PfxError("no GPU visible: the pfx kernels run on the device only")
    return torch


class Tables:
    """Device copy of a reaction network (pfx_tables_create)."""

    def __init__(self, tables: ReactionTables):
        lib = load_library()
        _torch()
        self.host = tables
        self._desc = tables.desc()
        h = C.c_void_p()
        rc = lib.pfx_tables_create(C.byref(self._desc), C.byref(h))
        if rc != 0:
            raise PfxError(f"pfx_tables_create failed: {rc}")
        self.handle = h

    def tune(self, kernel: int, variant: int = -1) -> int:
        """pfx_tables_tune: set (variant >= 0) or query the register-cap variant."""
        rc = load_library().pfx_tables_tune(self.handle, kernel, variant)
        if rc < 0:
            raise PfxError(f"pfx_tables_tune failed: {rc}")
        return rc if variant < 0 else variant

    def timing(self):
        """(per-cell kernel ms, assembly kernel ms) of the last call."""
        ms = np.zeros(2)
        load_library().pfx_tables_timing(self.handle, ms.ctypes.data)
        return float(ms[0]), float(ms[1])

    def close(self):
        if getattr(self, "handle", None):
            load_library().pfx_tables_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DeviceCells:
    """torch CUDA tensors for the pfx_cells fields (cell-major, FP64)."""
    t: Dict[str, object]
    ncells: int

    @classmethod
    def from_host(cls, arrays: Dict[str, Optional[np.ndarray]], ncells: int, device="cuda"):
        torch = _torch()
        t = {}
        for k, v in arrays.items():
            if v is None:
                t[k] = None
            else:
                t[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
        return cls(t, ncells)

    def struct(self) -> abi.Cells:
        c = abi.Cells()
        c.ncells = self.ncells
        for f in abi.CELL_FIELDS:
            v = self.t.get(f)
            setattr(c, f, None if v is None else v.data_ptr())
        return c

    def to_host(self) -> Dict[str, Optional[np.ndarray]]:
        return {k: (None if v is None else v.cpu().numpy()) for k, v in self.t.items()}

    def __getitem__(self, k):
        return self.t[k]


def react_batch(tables: Tables, cells: DeviceCells, opts: abi.ReactOpts):
    """pfx_react_batch: returns (rc, stats)."""
    lib = load_library()
    torch = _torch()
    cs = cells.struct()
    st = abi.Stats()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_react_batch(tables.handle, C.byref(cs), C.byref(opts), C.byref(st))
    if rc < 0:
        raise PfxError(f"pfx_react_batch failed: {rc}")
    return rc, st


def react_batch_noalloc(tables: Tables, cells: DeviceCells, opts: abi.ReactOpts):
    """pfx_react_batch_noalloc (no device allocation in the call): returns
    (rc, stats); rc < 0 is returned, not raised."""
    lib = load_library()
    torch = _torch()
    cs = cells.struct()
    st = abi.Stats()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    return lib.pfx_react_batch_noalloc(tables.handle, C.byref(cs), C.byref(opts), C.byref(st)), st


def residual_jacobian(tables: Tables, cells: DeviceCells, opts: abi.ReactOpts):
    """pfx_residual_jacobian: residual [n,N] and Jacobian [n,N,N] (row, col)."""
    lib = load_library()
    torch = _torch()
    n, N = cells.ncells, tables.host.N
    res = torch.zeros((n, N), dtype=torch.float64, device="cuda")
    jac = torch.zeros((n, N, N), dtype=torch.float64, device="cuda")  # [cell][col][row]
    cs = cells.struct()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_residual_jacobian(tables.handle, C.byref(cs), C.byref(opts), res.data_ptr(), jac.data_ptr())
    if rc != 0:
        raise PfxError(f"pfx_residual_jacobian failed: {rc}")
    torch.cuda.synchronize()
    return res.cpu().numpy(), jac.transpose(1, 2).contiguous().cpu().numpy()


def equilibrate(tables: Tables, cells: DeviceCells, constraint_type, constraint_conc, opts: abi.ReactOpts,
                molal: bool = False, rxn=None):
    """pfx_equilibrate_batch; constraint_conc is a [n, N] array (host or
    device); rxn a ctypes array [N] of abi.ConstraintRxn for mineral / gas
    rows (database.constraint_rxns)."""
    lib = load_library()
    torch = _torch()
    ct = np.ascontiguousarray(constraint_type, dtype=np.int32)
    cc = constraint_conc
    if not isinstance(cc, torch.Tensor):
        cc = torch.from_numpy(np.ascontiguousarray(cc, dtype=np.float64)).cuda()
    cs = cells.struct()
    st = abi.Stats()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_equilibrate_batch(tables.handle, C.byref(cs), ct.ctypes.data, cc.data_ptr(),
                                   None if rxn is None else C.byref(rxn), int(molal), C.byref(opts), C.byref(st))
    if rc < 0:
        raise PfxError(f"pfx_equilibrate_batch failed: {rc}")
    return rc, st


def auxvar(tables: Tables, cells: DeviceCells, activity_passes: int, opts: abi.ReactOpts):
    """pfx_auxvar_batch: totals / sorbed totals / complexes at the given free ions."""
    lib = load_library()
    torch = _torch()
    cs = cells.struct()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_auxvar_batch(tables.handle, C.byref(cs), int(activity_passes), C.byref(opts))
    if rc != 0:
        raise PfxError(f"pfx_auxvar_batch failed: {rc}")
    torch.cuda.synchronize()


def update_kinetic_state(tables: Tables, cells: DeviceCells, opts: abi.ReactOpts):
    """pfx_update_kinetic_state: mineral volume fractions after a converged step."""
    lib = load_library()
    torch = _torch()
    cs = cells.struct()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_update_kinetic_state(tables.handle, C.byref(cs), C.byref(opts))
    if rc != 0:
        raise PfxError(f"pfx_update_kinetic_state failed: {rc}")


class DeviceGrid:
    """Device copy of a structured grid (pfx_grid): face fluxes, dispersion
    coefficients and boundary data as torch CUDA tensors."""

    def __init__(self, nx, ny, nz, q, dd, area=(1.0, 1.0, 1.0), bc_q=(None, None), bc_dd=(None, None),
                 bc_total=(None, None), log_formulation=False):
        torch = _torch()
        dev = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)  # noqa
                                                               if not isinstance(a, torch.Tensor) else a,
                                                               dtype=torch.float64, device="cuda").contiguous()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.q = [dev(a) for a in q]
        self.dd = [dev(a) for a in dd]
        self.area = tuple(float(a) for a in area)
        self.bc_q = [dev(a) for a in bc_q]
        self.bc_dd = [dev(a) for a in bc_dd]
        self.bc_total = [dev(a) for a in bc_total]
        self.log_formulation = bool(log_formulation)

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    def struct(self) -> abi.Grid:
        g = abi.Grid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.log_formulation = int(self.log_formulation)
        ptr = lambda a: None if (a is None or a.numel() == 0) else a.data_ptr()  # noqa: E731
        for d in range(3):
            g.area[d] = self.area[d]
            g.q[d] = ptr(self.q[d])
            g.dd[d] = ptr(self.dd[d])
        for s_ in range(2):
            g.bc_q[s_] = ptr(self.bc_q[s_])
            g.bc_dd[s_] = ptr(self.bc_dd[s_])
            g.bc_total[s_] = ptr(self.bc_total[s_])
        return g


def gi_structure(grid: DeviceGrid):
    """pfx_gi_structure: (row_ptr, col_idx) int32 device tensors."""
    lib = load_library()
    torch = _torch()
    g = grid.struct()
    nnzb = C.c_int64(0)
    rc = lib.pfx_gi_structure(C.byref(g), None, None, C.byref(nnzb))
    if rc != 0:
        raise PfxError(f"pfx_gi_structure failed: {rc}")
    row_ptr = torch.empty(grid.ncells + 1, dtype=torch.int32, device="cuda")
    col_idx = torch.empty(nnzb.value, dtype=torch.int32, device="cuda")
    rc = lib.pfx_gi_structure(C.byref(g), row_ptr.data_ptr(), col_idx.data_ptr(), C.byref(nnzb))
    if rc != 0:
        raise PfxError(f"pfx_gi_structure failed: {rc}")
    return row_ptr, col_idx


def gi_residual_jacobian(tables: Tables, cells: DeviceCells, grid: DeviceGrid, opts: abi.ReactOpts, row_ptr,
                         residual=None, values=None):
    """pfx_gi_residual_jacobian: residual [n, N] and BSR values [nnzb, N, N]
    in memory order [block][col][row] (column-major blocks), device tensors."""
    lib = load_library()
    torch = _torch()
    N = tables.host.N
    nnzb = int(row_ptr[-1].item())
    if residual is None:
        residual = torch.empty((grid.ncells, N), dtype=torch.float64, device="cuda")
    if values is None:
        values = torch.empty((nnzb, N, N), dtype=torch.float64, device="cuda")
    cs = cells.struct()
    g = grid.struct()
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_gi_residual_jacobian(tables.handle, C.byref(cs), C.byref(g), C.byref(opts), row_ptr.data_ptr(),
                                      residual.data_ptr(), values.data_ptr())
    if rc != 0:
        raise PfxError(f"pfx_gi_residual_jacobian failed: {rc}")
    return residual, values


class Bsr:
    """pfx_bsr handle over a device BSR structure (row_ptr, col_idx int32
    tensors, kept alive here) with column-major N x N blocks."""

    def __init__(self, row_ptr, col_idx, N: int):
        self.lib = load_library()
        self.row_ptr, self.col_idx, self.N = row_ptr, col_idx, int(N)
        self.nrows = int(row_ptr.numel()) - 1
        h = C.c_void_p()
        rc = self.lib.pfx_bsr_create(self.nrows, self.N, int(col_idx.numel()), row_ptr.data_ptr(),
                                     col_idx.data_ptr(), C.byref(h))
        if rc != 0:
            raise PfxError(f"pfx_bsr_create failed: {rc}")
        self.handle = h

    def spmv(self, values, x, y=None):
        torch = _torch()
        if y is None:
            y = torch.empty((self.nrows, self.N), dtype=torch.float64, device="cuda")
        rc = self.lib.pfx_bsr_spmv(self.handle, values.data_ptr(), x.data_ptr(), y.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream)
        if rc != 0:
            raise PfxError(f"pfx_bsr_spmv failed: {rc}")
        return y

    def solve(self, values, b, x=None, rtol=1e-5, atol=0.0, dtol=0.0, max_iterations=0):
        """BiCGStab with point-block Jacobi from x = 0: (x, KrylovStats)."""
        torch = _torch()
        if x is None:
            x = torch.empty((self.nrows, self.N), dtype=torch.float64, device="cuda")
        o = abi.KrylovOpts(rtol=rtol, atol=atol, dtol=dtol, max_iterations=max_iterations)
        st = abi.KrylovStats()
        rc = self.lib.pfx_bsr_solve(self.handle, values.data_ptr(), b.data_ptr(), x.data_ptr(), C.byref(o),
                                    torch.cuda.current_stream().cuda_stream, C.byref(st))
        if rc != 0:
            raise PfxError(f"pfx_bsr_solve failed: {rc}")
        return x, st

    def close(self):
        if self.handle:
            self.lib.pfx_bsr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OsTransport:
    """pfx_os_transport_step: the implicit transport half of an
    operator-split step on a grid (the transport matrix with block size 1 on
    the grid's 7-point pattern, one BiCGStab solve per component)."""

    def __init__(self, grid: DeviceGrid):
        torch = _torch()
        self.lib = load_library()
        self.row_ptr, self.col_idx = gi_structure(grid)
        self.bsr = Bsr(self.row_ptr, self.col_idx, 1)
        n = grid.ncells
        self.work = torch.empty(int(self.col_idx.numel()) + 2 * n, dtype=torch.float64, device="cuda")

    def step(self, cells: DeviceCells, grid: DeviceGrid, ncomp: int, opts: abi.ReactOpts, rtol: float = 1e-12,
             max_iterations: int = 0):
        """Transport cells' totals by one step of opts.dt (in place):
        (return code 0 or 1 = some component did not converge, KrylovStats)."""
        torch = _torch()
        if opts.stream is None:
            opts.stream = torch.cuda.current_stream().cuda_stream
        ko = abi.KrylovOpts(rtol=rtol, max_iterations=max_iterations)
        st = abi.KrylovStats()
        cs = cells.struct()
        g = grid.struct()
        rc = self.lib.pfx_os_transport_step(self.bsr.handle, C.byref(g), C.byref(cs), int(ncomp), C.byref(opts),
                                            self.work.data_ptr(), C.byref(ko), C.byref(st))
        if rc < 0:
            raise PfxError(f"pfx_os_transport_step failed: {rc}")
        return rc, st

    def close(self):
        self.bsr.close()


class OsStepGraph:
    """pfx_os_graph_create / launch / stats: one operator-split step
    (transport of every component, reaction, volume-fraction update)
    captured as a HIP graph on `cells`, replayed with launch()."""

    def __init__(self, transport: OsTransport, grid: DeviceGrid, tables: Tables, cells: DeviceCells, ncomp: int,
                 transport_opts: abi.ReactOpts, react_opts: abi.ReactOpts, rtol: float = 1e-12,
                 max_iterations: int = 60):
        self.lib = load_library()
        self._keep = (transport, grid, tables, cells, grid.struct(), cells.struct())
        ko = abi.KrylovOpts(rtol=rtol, max_iterations=max_iterations)
        h = C.c_void_p()
        rc = self.lib.pfx_os_graph_create(transport.bsr.handle, C.byref(self._keep[4]), tables.handle,
                                          C.byref(self._keep[5]), int(ncomp), C.byref(transport_opts),
                                          C.byref(react_opts), transport.work.data_ptr(), C.byref(ko), C.byref(h))
        if rc != 0:
            raise PfxError(f"pfx_os_graph_create failed: {rc}")
        self.handle = h

    def launch(self):
        torch = _torch()
        rc = self.lib.pfx_os_graph_launch(self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc != 0:
            raise PfxError(f"pfx_os_graph_launch failed: {rc}")

    def stats(self):
        """(reaction return code, reaction Stats, transport KrylovStats) of the
        last replay."""
        rc = C.c_int32()
        st, ks = abi.Stats(), abi.KrylovStats()
        r = self.lib.pfx_os_graph_stats(self.handle, C.byref(rc), C.byref(st), C.byref(ks))
        if r != 0:
            raise PfxError(f"pfx_os_graph_stats failed: {r}")
        return rc.value, st, ks

    def close(self):
        if getattr(self, "handle", None):
            self.lib.pfx_os_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GiSolver:
    """pfx_gi_solver: one global-implicit time step (Newton + BiCGStab) on
    a grid of the tables' network."""

    def __init__(self, tables: Tables, grid: DeviceGrid):
        self.lib = load_library()
        self.tables = tables
        h = C.c_void_p()
        g = grid.struct()
        rc = self.lib.pfx_gi_solver_create(tables.handle, C.byref(g), C.byref(h))
        if rc != 0:
            raise PfxError(f"pfx_gi_solver_create failed: {rc}")
        self.handle = h

    def step(self, cells: DeviceCells, grid: DeviceGrid, opts: abi.ReactOpts, newton: Optional[abi.NewtonOpts] = None):
        torch = _torch()
        if opts.stream is None:
            opts.stream = torch.cuda.current_stream().cuda_stream
        no = newton if newton is not None else abi.NewtonOpts()
        st = abi.NewtonStats()
        cs = cells.struct()
        g = grid.struct()
        rc = self.lib.pfx_gi_solver_step(self.handle, C.byref(cs), C.byref(g), C.byref(opts), C.byref(no),
                                         C.byref(st))
        if rc != 0:
            raise PfxError(f"pfx_gi_solver_step failed: {rc}")
        return st

    def timestep(self, cells: DeviceCells, grid: DeviceGrid, opts: abi.ReactOpts,
                 newton: Optional[abi.NewtonOpts] = None, max_cuts: int = -1, reduction_factor: float = 0.5):
        """pfx_gi_solver_timestep: the step with the time-step cut and
        retry; returns abi.GiStepStats (dt = the step taken)."""
        torch = _torch()
        if opts.stream is None:
            opts.stream = torch.cuda.current_stream().cuda_stream
        no = newton if newton is not None else abi.NewtonOpts()
        st = abi.GiStepStats()
        cs = cells.struct()
        g = grid.struct()
        rc = self.lib.pfx_gi_solver_timestep(self.handle, C.byref(cs), C.byref(g), C.byref(opts), C.byref(no),
                                             C.c_int32(max_cuts), C.c_double(reduction_factor), C.byref(st))
        if rc != 0:
            raise PfxError(f"pfx_gi_solver_timestep failed: {rc}")
        return st

    def close(self):
        if self.handle:
            self.lib.pfx_gi_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gi_initial_iterate(tables: Tables, cells: DeviceCells, opts: abi.ReactOpts, warm=None):
    """The global-implicit step's starting iterate from each cell's own
    reaction step.  timestepper_BE starts SNES from the old solution
    (timestepper_BE.F90:326); here every cell first solves its reaction step
    alone (pfx_react_batch: the cell's accumulation and rates over opts.dt,
    no transport) on a copy of the cells, and where that converged its free
    ions and exchange state become the Newton's starting iterate in `cells`.
    cells.total / total_sorb (the old accumulation) and the activity
    coefficients are left as they are, so the global residual, Jacobian and
    solution are those of the step from the old state; only the path of the
    global Newton changes.  warm(tables, copy) may fill copy['guess'] (a
    learned warm start, solved with PFX_GUESS_ARRAY_FALLBACK from the old free
    ions as fallback); without it the reaction step starts from the old free
    ions.  Returns the reaction step's statistics."""
    torch = _torch()
    work = DeviceCells({k: (None if v is None else v.clone()) for k, v in cells.t.items()}, cells.ncells)
    o = abi.ReactOpts()
    C.memmove(C.addressof(o), C.addressof(opts), C.sizeof(o))
    if warm is not None:
        if work.t.get("guess") is None:
            work.t["guess"] = torch.empty_like(work.t["pri_molal"])
        warm(tables, work)
        o.guess_mode = abi.PFX_GUESS_ARRAY_FALLBACK
        o.fallback_previous = 1
    else:
        o.guess_mode = abi.PFX_GUESS_PREVIOUS
    _, st = react_batch(tables, work, o)
    ok = (work.t["status"] == 0).unsqueeze(1)
    cells.t["pri_molal"].copy_(torch.where(ok, work.t["pri_molal"], cells.t["pri_molal"]))
    if cells.t.get("ionx_ref_sorbed") is not None and work.t.get("ionx_ref_sorbed") is not None:
        cells.t["ionx_ref_sorbed"].copy_(torch.where(ok, work.t["ionx_ref_sorbed"], cells.t["ionx_ref_sorbed"]))
    del work
    return st


def gi_dispersion(grid: DeviceGrid, porosity, saturation=None, tortuosity=None, dispersivity=0.0, diffusion=1e-9,
                  dx=(1.0, 1.0, 1.0), bc_type=(0, 0)):
    """pfx_gi_dispersion: fills grid.dd and grid.bc_dd (device) from cell
    properties (device tensors [ncells]) and the grid's face fluxes."""
    lib = load_library()
    torch = _torch()
    for d in range(3):
        if grid.dd[d] is None or grid.dd[d].shape != grid.q[d].shape:
            grid.dd[d] = torch.zeros_like(grid.q[d])
    for s_ in range(2):
        if grid.bc_dd[s_] is None:
            grid.bc_dd[s_] = torch.zeros((grid.nz, grid.ny), dtype=torch.float64, device="cuda")
    g = grid.struct()
    ptr = lambda a: None if a is None else a.data_ptr()  # noqa: E731
    ddp = (C.c_void_p * 3)(*[ptr(a) if a.numel() else None for a in grid.dd])
    bcp = (C.c_void_p * 2)(*[ptr(a) for a in grid.bc_dd])
    dxa = np.ascontiguousarray(dx, dtype=np.float64)
    bt = np.ascontiguousarray(bc_type, dtype=np.int32)
    rc = lib.pfx_gi_dispersion(C.byref(g), ptr(porosity), ptr(saturation), ptr(tortuosity), dispersivity, diffusion,
                               dxa.ctypes.data, bt.ctypes.data, ddp, bcp,
                               C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise PfxError(f"pfx_gi_dispersion failed: {rc}")


def _zone_arrays(tables, cells):
    nz = len(tables)
    if nz != len(cells) or not 1 <= nz <= abi.PFX_MAX_ZONES:
        raise PfxError("one cell batch per zone, at most PFX_MAX_ZONES zones")
    handles = (C.c_void_p * nz)(*[t.handle.value for t in tables])
    structs = (abi.Cells * nz)(*[c.struct() for c in cells])
    return nz, handles, structs


def react_zones(tables, cells, opts: abi.ReactOpts, plan: Optional[abi.ZonePlan] = None):
    """pfx_react_zones: one persistent launch over several chemistry zones.
    Returns (rc, [stats per zone], plan)."""
    lib = load_library()
    torch = _torch()
    nz, handles, structs = _zone_arrays(tables, cells)
    stats = (abi.Stats * nz)()
    if plan is None:
        plan = abi.ZonePlan()
        plan.nzones = nz
    if opts.stream is None:
        opts.stream = torch.cuda.current_stream().cuda_stream
    rc = lib.pfx_react_zones(nz, handles, structs, C.byref(opts), C.byref(plan), stats)
    if rc < 0:
        raise PfxError(f"pfx_react_zones failed: {rc}")
    return rc, list(stats), plan


def rebalance(plan: abi.ZonePlan):
    rc = load_library().pfx_zone_plan_rebalance(C.byref(plan))
    if rc != 0:
        raise PfxError(f"pfx_zone_plan_rebalance failed: {rc}")
    return plan


class ZonesGraph:
    """pfx_zones_graph_create / launch / stats: a captured zoned step."""

    def __init__(self, tables, cells, opts: abi.ReactOpts, plan: Optional[abi.ZonePlan] = None):
        lib = load_library()
        _torch()
        nz, handles, structs = _zone_arrays(tables, cells)
        self.nz = nz
        self._keep = (tables, cells, structs)
        h = C.c_void_p()
        rc = lib.pfx_zones_graph_create(nz, handles, structs, C.byref(opts), C.byref(plan) if plan else None,
                                        C.byref(h))
        if rc != 0:
            raise PfxError(f"pfx_zones_graph_create failed: {rc}")
        self.handle = h

    def launch(self):
        import torch
        rc = load_library().pfx_graph_launch(self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc != 0:
            raise PfxError(f"pfx_graph_launch failed: {rc}")

    def stats(self):
        stats = (abi.Stats * self.nz)()
        plan = abi.ZonePlan()
        rc = load_library().pfx_graph_stats(self.handle, stats, C.byref(plan))
        if rc < 0:
            raise PfxError(f"pfx_graph_stats failed: {rc}")
        return rc, list(stats), plan

    def close(self):
        if getattr(self, "handle", None):
            load_library().pfx_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

