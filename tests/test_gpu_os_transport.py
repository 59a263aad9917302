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
