"""The operator-split transport half on the device (pfx_os_transport_step)
against the oracle's direct solve of the same matrix
(oracle.os_transport_step, pinned in tests/test_os_transport.py), and the
device-resident operator-split column (implicit transport, reaction,
kinetic-state update) against the same steps run by the oracle."""
import numpy as np
import pytest

import oracle as orc
from chemdefs import carbonate_column
from pflotran_amd import abi, column, runtime, synth
from test_column_os import initial_state, react_opts

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def random_grid(rng, nx, ny, nz, N, dirichlet=True):
    shapes = [(nz, ny, nx - 1), (nz, ny - 1, nx), (nz - 1, ny, nx)]
    q = [rng.normal(0, 1e-6, size=s) for s in shapes]
    dd = [rng.uniform(0.2, 1.0, size=s) * 1e-7 for s in shapes]
    bc_q = (np.full((nz, ny), 3e-6), np.full((nz, ny), -3e-6))
    bc_dd = (np.full((nz, ny), 2e-7), np.zeros((nz, ny)))
    bc_total = (rng.uniform(1e-4, 1e-2, size=N) if dirichlet else None, None)
    return dict(q=q, dd=dd, area=(1.0, 2.0, 0.5), bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total)


@pytest.mark.parametrize("shape,N", [((6, 5, 4), 4), ((100, 1, 1), 5), ((16, 16, 8), 15)])
def test_transport_step_matches_direct_solve(shape, N):
    import torch
    rng = np.random.default_rng(sum(shape) + N)
    nx, ny, nz = shape
    n = nx * ny * nz
    f = random_grid(rng, nx, ny, nz, N)
    phi = rng.uniform(0.1, 0.4, size=n)
    vol = rng.uniform(0.5, 2.0, size=n)
    dt = 3600.0
    T = rng.uniform(1e-5, 1e-2, size=(n, N))
    hg = orc.HostGrid(nx, ny, nz, **f)
    acc = phi * 1.0 * 1000.0 * vol / dt
    ref = orc.os_transport_step(hg, acc, T)
    g = runtime.DeviceGrid(nx, ny, nz, **f)
    dc = runtime.DeviceCells({"total": torch.as_tensor(T, device="cuda"), "porosity": torch.as_tensor(phi, device="cuda"),
                              "volume": torch.as_tensor(vol, device="cuda")}, n)
    tr = runtime.OsTransport(g)
    print(f"transport step {shape} N={N}", flush=True)
    rc, st = tr.step(dc, g, N, abi.default_opts(dt=dt, saturation=1.0), rtol=1e-13)
    assert rc == 0 and st.reason == abi.PFX_KSP_CONVERGED_RTOL, (st.reason, st.iterations)
    out = dc.t["total"].cpu().numpy()
    assert rel_err(out, ref) <= 1e-10
    tr.close()


def test_closed_box_conserves_on_device():
    import torch
    rng = np.random.default_rng(9)
    nx, ny, nz, N = 8, 6, 5, 3
    f = random_grid(rng, nx, ny, nz, N)
    f.update(bc_q=(None, None), bc_dd=(None, None), bc_total=(None, None))
    n = nx * ny * nz
    T = rng.uniform(1e-5, 1e-2, size=(n, N))
    g = runtime.DeviceGrid(nx, ny, nz, **f)
    dc = runtime.DeviceCells({"total": torch.as_tensor(T, device="cuda")}, n)
    tr = runtime.OsTransport(g)
    o = abi.default_opts(dt=7200.0, porosity=0.3, saturation=1.0, volume=1.0)
    for _ in range(3):
        rc, st = tr.step(dc, g, N, o, rtol=1e-13)
        assert rc == 0
    out = dc.t["total"].cpu().numpy()
    assert np.allclose(out.sum(0), T.sum(0), rtol=1e-11, atol=0)
    assert not np.allclose(out, T)
    tr.close()


def test_device_resident_operator_split_column():
    """The 1-D carbonate column with the implicit transport half: every
    step on the device, against the oracle's implicit transport (direct
    solve) plus its batched reaction step."""
    t = carbonate_column()
    col = column.Column()
    cells, inflow = initial_state(t, col.ncells)
    ref = orc.HostCells({k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}, col.ncells)
    nsteps, dt = 10, 86400.0
    f = column.grid_faces(col, inflow)
    hg = orc.HostGrid(f["nx"], f["ny"], f["nz"], f["q"], f["dd"], area=f["area"], bc_q=f["bc_q"], bc_dd=f["bc_dd"],
                      bc_total=f["bc_total"])
    acc = np.full(col.ncells, col.porosity * col.saturation * 1000.0 * col.volume / dt)
    for _ in range(nsteps):
        ref.arrays["total"] = np.ascontiguousarray(orc.os_transport_step(hg, acc, ref["total"]))
        rc, st = orc.react_batch(t, ref, react_opts(dt))
        assert rc == 0
        orc.update_kinetic_state(t, ref, react_opts(dt))
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host({k: v for k, v in cells.arrays.items()}, col.ncells)
    print("device-resident column: 10 steps", flush=True)
    stats, tstats = column.os_steps_device_implicit(tab, dc, col, inflow, [dt] * nsteps, react_opts)
    assert all(s.converged == col.ncells for s in stats)
    assert all(ts.reason == abi.PFX_KSP_CONVERGED_RTOL for ts in tstats)
    out = dc.to_host()
    assert rel_err(out["pri_molal"], ref["pri_molal"]) <= 1e-8
    assert rel_err(out["total"], ref["total"]) <= 1e-8
    assert np.max(np.abs(out["mnrl_volfrac"] - ref["mnrl_volfrac"])) <= 1e-12
    # the inflow water has moved into the column
    assert rel_err(out["total"][0], inflow) < rel_err(cells["total"][0], inflow)


def test_multi_rhs_solve_matches_single_solves():
    # pfx_bsr_solve_multi: every column solved by the same BiCGStab as
    # pfx_bsr_solve (same scalars and tests per column; the dot products are
    # summed in another order, so iterates agree to rounding), against the
    # direct solve; stream-ordered, it gives the same result
    import ctypes as C

    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    rng = np.random.default_rng(3)
    nx, ny, nz, N = 9, 7, 5, 6
    n = nx * ny * nz
    f = random_grid(rng, nx, ny, nz, N)
    g = runtime.DeviceGrid(nx, ny, nz, **f)
    rp, ci = runtime.gi_structure(g)
    rp_h, ci_h = rp.cpu().numpy(), ci.cpu().numpy()
    vals = rng.normal(size=len(ci_h))
    for row in range(n):
        for k in range(rp_h[row], rp_h[row + 1]):
            if ci_h[k] == row:
                vals[k] = 8.0 + abs(vals[k])
    A = sp.csr_matrix((vals, ci_h, rp_h), shape=(n, n))
    Bm = rng.normal(size=(n, N))
    Bm[:, 2] = 0.0                        # a zero right-hand side: decided at the start
    want = np.stack([spla.spsolve(A.tocsc(), Bm[:, c]) for c in range(N)], axis=1)
    bsr = runtime.Bsr(rp, ci, 1)
    lib = runtime.load_library()
    dv = torch.as_tensor(vals, device="cuda")
    dB = torch.as_tensor(Bm, device="cuda")
    for ordered in (0, 1):
        X = torch.zeros_like(dB)
        ko = abi.KrylovOpts(rtol=1e-13, max_iterations=200)
        st = abi.KrylovStats()
        s = torch.cuda.current_stream().cuda_stream
        rc = lib.pfx_bsr_solve_multi(bsr.handle, dv.data_ptr(), dB.data_ptr(), X.data_ptr(), N, C.byref(ko),
                                     C.c_void_p(s), ordered, C.byref(st))
        assert rc == 0
        if ordered:
            assert lib.pfx_bsr_multi_stats(bsr.handle, C.c_void_p(s), C.byref(st)) == 0
        # the zero column decides at the start on atol (rtol |b| = 0): the
        # worst reason of the columns is an atol convergence
        assert st.reason == abi.PFX_KSP_CONVERGED_ATOL, (ordered, st.reason)
        x = X.cpu().numpy()
        assert np.all(x[:, 2] == 0.0)
        assert np.max(np.abs(x - want)) <= 1e-10 * np.max(np.abs(want))
        # each column against its own single-right-hand-side solve
        for c in (0, 4):
            xs = torch.zeros(n, dtype=torch.float64, device="cuda")
            rs, ss = bsr.solve(dv, dB[:, c].contiguous(), xs, rtol=1e-13, max_iterations=200)
            assert np.max(np.abs(xs.cpu().numpy() - x[:, c])) <= 1e-12 * np.max(np.abs(want[:, c]))
    bsr.close()


def test_operator_split_step_graph_replays_match_eager():
    """One whole operator-split step (transport of every component,
    reaction, volume fractions) captured as a graph: each replay, with the
    eager steps and synchronisations in between, leaves the same bits as the
    eager step on a twin of the cells."""
    import torch
    t = carbonate_column()
    col = column.Column()
    cells, inflow = initial_state(t, col.ncells)
    dt = 86400.0
    N = t.N
    tab = runtime.Tables(t)
    f = column.grid_faces(col, inflow)
    grid = runtime.DeviceGrid(f["nx"], f["ny"], f["nz"], f["q"], f["dd"], area=f["area"], bc_q=f["bc_q"],
                              bc_dd=f["bc_dd"], bc_total=f["bc_total"])
    o_t = abi.default_opts(dt=dt, porosity=col.porosity, saturation=col.saturation, volume=col.volume)
    de = runtime.DeviceCells.from_host({k: v for k, v in cells.arrays.items()}, col.ncells)
    dg = runtime.DeviceCells.from_host({k: v for k, v in cells.arrays.items()}, col.ncells)
    tr_e, tr_g = runtime.OsTransport(grid), runtime.OsTransport(grid)
    # a warm-up step on a scratch twin sizes the library's buffers before capture
    dw = runtime.DeviceCells.from_host({k: v for k, v in cells.arrays.items()}, col.ncells)
    tr_g.step(dw, grid, N, o_t, rtol=1e-12, max_iterations=60)
    runtime.react_batch(tab, dw, react_opts(dt))
    print("operator-split step graph: capture", flush=True)
    graph = runtime.OsStepGraph(tr_g, grid, tab, dg, N, o_t, react_opts(dt), rtol=1e-12, max_iterations=60)
    for step in range(5):
        rc, ts = tr_e.step(de, grid, N, o_t, rtol=1e-12, max_iterations=60)
        assert rc == 0
        rc, st = runtime.react_batch(tab, de, react_opts(dt))
        runtime.update_kinetic_state(tab, de, react_opts(dt))
        torch.cuda.synchronize()
        graph.launch()
        rc_g, st_g, ks_g = graph.stats()
        print(f"replay {step}: transport its {ks_g.iterations} (eager {ts.iterations})", flush=True)
        assert rc_g == rc == 0 and st_g.converged == st.converged == col.ncells
        assert ks_g.reason == ts.reason == abi.PFX_KSP_CONVERGED_RTOL and ks_g.iterations == ts.iterations
        for k in ("total", "pri_molal", "mnrl_volfrac", "pri_act_coef"):
            assert torch.equal(dg.t[k], de.t[k]), (step, k)
    graph.close()
    tr_e.close()
    tr_g.close()
