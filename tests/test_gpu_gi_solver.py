"""The global-implicit step on the device (include/pfx_gi_solver.h) against
what the mathematics fixes: the block product against scipy's BSR product,
the preconditioned BiCGStab solution against scipy's direct solve (and its
true residual), and each Newton step against the Newton step whose linear
system is solved directly (scipy) on the oracle's residual and Jacobian.
The Krylov iterates themselves are PETSc internals with no reference values
(parity unpinned); the Newton iterates are pinned through the linear
systems they solve.  A one-cell grid with closed faces is the per-cell
reaction system, so its step must land on pfx_react_batch's solution."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as orc
from gpu_cases import host_arrays, speciated_state
from pflotran_amd import abi, runtime, synth

pytestmark = pytest.mark.gpu


def faces(rng, nx, ny, nz):
    shapes = [(nz, ny, nx - 1), (nz, ny - 1, nx), (nz - 1, ny, nx)]
    q = [rng.normal(0, 1e-6, size=s) for s in shapes]
    dd = [rng.uniform(0.2, 1.0, size=s) * 1e-7 for s in shapes]
    return q, dd


def scipy_bsr(vals, row_ptr, col_idx, N, colmajor=True):
    """scipy BSR of blocks stored [block][col][row] (the library's
    column-major blocks) or, colmajor=False, [block][row][col]."""
    n = len(row_ptr) - 1
    blocks = np.ascontiguousarray(vals.transpose(0, 2, 1)) if colmajor else vals
    return sp.bsr_matrix((blocks, col_idx, row_ptr), shape=(n * N, n * N))


def random_system(rng, nx, ny, nz, N, dominance=3.0):
    import torch
    g = runtime.DeviceGrid(nx, ny, nz, *faces(rng, nx, ny, nz))
    rp, ci = runtime.gi_structure(g)
    rp_h, ci_h = rp.cpu().numpy(), ci.cpu().numpy()
    vals = rng.normal(size=(len(ci_h), N, N))
    for row in range(len(rp_h) - 1):
        for b in range(rp_h[row], rp_h[row + 1]):
            if ci_h[b] == row:
                vals[b] += dominance * N * np.eye(N)
    return rp, ci, rp_h, ci_h, vals, torch.as_tensor(vals, device="cuda")


@pytest.mark.parametrize("N", [15, 8, 30])
def test_bsr_product_and_solve_against_scipy(N):
    import torch
    rng = np.random.default_rng(N)
    rp, ci, rp_h, ci_h, vals, dvals = random_system(rng, 6, 5, 4, N)
    A = scipy_bsr(vals, rp_h, ci_h, N)
    n = len(rp_h) - 1
    bsr = runtime.Bsr(rp, ci, N)
    x = rng.normal(size=(n, N))
    y = bsr.spmv(dvals, torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = A @ x.ravel()
    assert np.max(np.abs(y.ravel() - ref)) <= 1e-13 * np.max(np.abs(ref))
    b = rng.normal(size=(n, N))
    xs, st = bsr.solve(dvals, torch.as_tensor(b, device="cuda"), rtol=1e-12)
    xs = xs.cpu().numpy().ravel()
    assert st.reason == abi.PFX_KSP_CONVERGED_RTOL and 0 < st.iterations < 200
    true_res = np.linalg.norm(A @ xs - b.ravel())
    assert true_res <= 1e-11 * np.linalg.norm(b)
    xd = spla.spsolve(A.tocsc(), b.ravel())
    assert np.max(np.abs(xs - xd)) <= 1e-9 * np.max(np.abs(xd))
    assert abs(st.rnorm0 - np.linalg.norm(b)) <= 1e-12 * np.linalg.norm(b)
    # the default (PETSc) tolerance stops early, at rtol 1e-5 of ||b||
    xl, st2 = bsr.solve(dvals, torch.as_tensor(b, device="cuda"))
    assert st2.reason == abi.PFX_KSP_CONVERGED_RTOL and st2.iterations <= st.iterations
    assert np.linalg.norm(A @ xl.cpu().numpy().ravel() - b.ravel()) <= 2e-5 * np.linalg.norm(b)
    bsr.close()


def test_singular_diagonal_block_reported():
    import torch
    rng = np.random.default_rng(3)
    N = 15
    rp, ci, rp_h, ci_h, vals, dvals = random_system(rng, 3, 3, 2, N)
    diag0 = next(b for b in range(rp_h[0], rp_h[1]) if ci_h[b] == 0)
    dvals[diag0] = 0.0
    bsr = runtime.Bsr(rp, ci, N)
    _, st = bsr.solve(dvals, torch.ones((len(rp_h) - 1, N), dtype=torch.float64, device="cuda"))
    assert st.reason == abi.PFX_KSP_DIVERGED_PC_FAILED and st.iterations == 0
    bsr.close()


def cpu_newton(t, cells, hg, opts, N, max_it, rtol, stol=1e-8, atol=1e-50, max_dlnC=5.0):
    """Newton's method in SNESConvergedDefault's order with each linear system
    J y = F solved directly (scipy) on the oracle's residual and Jacobian,
    and PMRTCheckUpdatePre's clip of the log-form step."""
    history = []
    r, v, rp, ci = orc.gi_residual_jacobian(t, cells, hg, opts)
    f0 = np.linalg.norm(r)
    history.append(f0)
    reason, it = 0, 0
    while not reason:
        if it >= max_it:
            reason = abi.PFX_SNES_DIVERGED_MAX_IT
            break
        y = spla.spsolve(scipy_bsr(v, rp, ci, N, colmajor=False).tocsc(), r.ravel())
        y = np.sign(y) * np.minimum(np.abs(y), max_dlnC)
        m = cells["pri_molal"].ravel() * np.exp(-y)
        cells["pri_molal"][:] = m.reshape(cells["pri_molal"].shape)
        it += 1
        r, v, rp, ci = orc.gi_residual_jacobian(t, cells, hg, opts)
        f = np.linalg.norm(r)
        history.append(f)
        if f < atol:
            reason = abi.PFX_SNES_CONVERGED_FNORM_ABS
        elif f <= rtol * f0:
            reason = abi.PFX_SNES_CONVERGED_FNORM_RELATIVE
        elif np.linalg.norm(y) < stol * np.linalg.norm(np.log(m)):
            reason = abi.PFX_SNES_CONVERGED_SNORM_RELATIVE
    return it, reason, history


def test_gi_newton_step_matches_direct_solves():
    t = synth.random_network(15, 50, 3, ionx=True)
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    rng = np.random.default_rng(41)
    q, dd = faces(rng, nx, ny, nz)
    # inflow through a Dirichlet x-min face, outflow through a zero-gradient
    # x-max face without dispersion: the reference leaves the boundary
    # state's derivative out of a zero-gradient face's Jacobian, which is
    # exact only when that face carries no inflow and no dispersion
    bc_q = (np.full((nz, ny), 2e-5), np.full((nz, ny), -2e-5))
    bc_dd = (np.full((nz, ny), 2e-7), np.zeros((nz, ny)))
    cells, _, _, _ = speciated_state(t, n, seed=43)
    bc_total = (cells["total"][0] * 3.0, None)
    hg = orc.HostGrid(nx, ny, nz, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    opts = synth.react_opts()
    ref = cells.copy()
    its_c, reason_c, hist_c = cpu_newton(t, ref, hg, opts, t.N, max_it=20, rtol=1e-10)
    assert reason_c > 0 and its_c >= 2

    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    g = runtime.DeviceGrid(nx, ny, nz, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    gs = runtime.GiSolver(tab, g)
    no = abi.NewtonOpts(rtol=1e-10, max_iterations=20)
    no.krylov.rtol = 1e-13
    st = gs.step(dc, g, synth.react_opts(), no)
    out = dc.to_host()
    # the same linear systems up to rounding: the same path (the clipped
    # first steps, then quadratic convergence)
    assert st.reason == reason_c and abs(st.iterations - its_c) <= 1, (st.as_dict(), its_c, reason_c)
    assert abs(st.fnorm0 - hist_c[0]) <= 1e-10 * hist_c[0]
    assert st.fnorm <= 1e-10 * st.fnorm0
    assert np.max(np.abs(out["pri_molal"] - ref["pri_molal"]) / ref["pri_molal"]) <= 1e-9
    assert st.linear_iterations > 0 and st.ms_assembly > 0 and st.ms_linear > 0
    gs.close()


def test_predicted_iterate_reaches_the_same_solution():
    # runtime.gi_initial_iterate: each cell's reaction step first, then the
    # global Newton from its free ions.  The residual, Jacobian and tests are
    # those of the step from the old state, so both runs land on the oracle's
    # Newton solution; the predicted start needs no more Newton iterations.
    # A learned warm start is stood in for by the old free ions scaled by
    # 1.02 (it only seeds the reaction step, which falls back to the old
    # free ions where it fails).
    t = synth.random_network(15, 50, 3, ionx=True)
    nx, ny, nz = 5, 4, 3
    n = nx * ny * nz
    rng = np.random.default_rng(41)
    q, dd = faces(rng, nx, ny, nz)
    bc_q = (np.full((nz, ny), 2e-5), np.full((nz, ny), -2e-5))
    bc_dd = (np.full((nz, ny), 2e-7), np.zeros((nz, ny)))
    cells, _, _, _ = speciated_state(t, n, seed=43)
    bc_total = (cells["total"][0] * 3.0, None)
    hg = orc.HostGrid(nx, ny, nz, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    ref = cells.copy()
    its_c, reason_c, _ = cpu_newton(t, ref, hg, synth.react_opts(), t.N, max_it=20, rtol=1e-10)
    assert reason_c > 0
    tab = runtime.Tables(t)
    g = runtime.DeviceGrid(nx, ny, nz, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    no = abi.NewtonOpts(rtol=1e-10, max_iterations=20)
    no.krylov.rtol = 1e-13

    def warm(tb, c):
        c.t["guess"].copy_(c.t["pri_molal"] * 1.02)

    for how in ("old", "previous", "warm"):
        dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
        old_total = dc.t["total"].clone()
        if how != "old":
            pst = runtime.gi_initial_iterate(tab, dc, synth.react_opts(), warm=warm if how == "warm" else None)
            assert pst.converged == n
            # the old accumulation is untouched; the iterate moved
            assert bool((dc.t["total"] == old_total).all())
        gs = runtime.GiSolver(tab, g)
        st = gs.step(dc, g, synth.react_opts(), no)
        gs.close()
        out = dc.to_host()
        assert st.reason > 0, (how, st.as_dict())
        assert st.iterations <= its_c + 1, (how, st.iterations, its_c)
        assert np.max(np.abs(out["pri_molal"] - ref["pri_molal"]) / ref["pri_molal"]) <= 1e-9, how


def test_one_closed_cell_is_the_reaction_step():
    # no faces, no boundaries: the global-implicit system of one cell is the
    # dt-consistent per-cell system pfx_react_batch solves
    t = synth.random_network(15, 50, 3, ionx=True, act_freq=abi.PFX_ACT_OFF)
    cells, _, _, _ = speciated_state(t, 1, seed=47)
    tab = runtime.Tables(t)
    dr = runtime.DeviceCells.from_host(host_arrays(cells), 1)
    runtime.react_batch(tab, dr, synth.react_opts(extra_iterations=2, guess_mode=abi.PFX_GUESS_PREVIOUS))
    m_react = dr.to_host()["pri_molal"]
    empty = [np.zeros((1, 1, 0)), np.zeros((1, 0, 1)), np.zeros((0, 1, 1))]
    g = runtime.DeviceGrid(1, 1, 1, empty, empty, log_formulation=True)
    dg = runtime.DeviceCells.from_host(host_arrays(cells), 1)
    gs = runtime.GiSolver(tab, g)
    no = abi.NewtonOpts(rtol=1e-9, max_iterations=40)
    no.krylov.rtol = 1e-14
    st = gs.step(dg, g, synth.react_opts(), no)
    assert st.reason in (abi.PFX_SNES_CONVERGED_FNORM_RELATIVE, abi.PFX_SNES_CONVERGED_SNORM_RELATIVE,
                         abi.PFX_SNES_CONVERGED_FNORM_ABS), st.as_dict()
    m_gi = dg.to_host()["pri_molal"]
    assert np.max(np.abs(m_gi - m_react) / m_react) <= 1e-9
    gs.close()


def test_time_step_cut_restores_the_old_state_and_retries():
    # a column of identical cells fed through a Dirichlet inflow face with
    # 30x their totals (no minerals or exchange: the Newton count falls with
    # dt, 18 iterations at 1e6 s down to 3); a budget of 5 iterations fails
    # at the full step, which is cut by halves (timestepper_BE.F90:344-399)
    # until 5 suffice, each retry from the restored old state, so the
    # accepted attempt is the plain Newton solve of the old state at the
    # cut step
    t = synth.random_network(15, 50, 0, ionx=False)
    nx = 6
    one, _, _, _ = speciated_state(t, 1, seed=59)
    cells = orc.HostCells({k: (None if v is None else np.repeat(v, nx, axis=0)) for k, v in one.arrays.items()}, nx)
    orc.auxvar(t, cells, synth.react_opts(), update_activity=False)
    e = [np.zeros((1, 0, nx)), np.zeros((0, 1, nx))]
    q = [np.full((1, 1, nx - 1), 1e-6)] + e
    dd = [np.full((1, 1, nx - 1), 1e-9)] + e
    bc_q = (np.full((1, 1), 1e-6), np.full((1, 1), -1e-6))
    bc_dd = (np.full((1, 1), 2e-9), np.zeros((1, 1)))
    bc_total = (cells["total"][0] * 30.0, None)
    tab = runtime.Tables(t)
    g = runtime.DeviceGrid(nx, 1, 1, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    gs = runtime.GiSolver(tab, g)
    no = abi.NewtonOpts(rtol=1e-8, max_iterations=5)
    no.krylov.rtol = 1e-13
    dc = runtime.DeviceCells.from_host(host_arrays(cells), nx)
    dt0 = 1e6
    st = gs.timestep(dc, g, synth.react_opts(dt=dt0), no)
    assert st.reason > 0 and 3 <= st.cuts <= 12, st.as_dict()
    assert st.dt == dt0 * 0.5 ** st.cuts
    assert st.newton_iterations == st.wasted_newton_iterations + st.newton.iterations
    assert st.wasted_newton_iterations == 5 * st.cuts
    out = dc.to_host()
    ref = runtime.DeviceCells.from_host(host_arrays(cells), nx)
    st1 = gs.step(ref, g, synth.react_opts(dt=st.dt), no)
    assert st1.reason == st.reason and st1.iterations == st.newton.iterations
    assert np.array_equal(ref.to_host()["pri_molal"], out["pri_molal"])
    # the same step on the CPU (oracle residual, direct solves) needs more
    # than 5 iterations at the step before the last cut and at most 5 at it
    hg = orc.HostGrid(nx, 1, 1, q, dd, bc_q=bc_q, bc_dd=bc_dd, bc_total=bc_total, log_formulation=True)
    its_cut, r_cut, _ = cpu_newton(t, cells.copy(), hg, synth.react_opts(dt=st.dt), t.N, max_it=50, rtol=1e-8)
    its_prev, _, _ = cpu_newton(t, cells.copy(), hg, synth.react_opts(dt=2 * st.dt), t.N, max_it=50, rtol=1e-8)
    assert r_cut > 0 and its_cut <= 5 < its_prev
    # a cut budget of 0 leaves the failed step in place and reports it
    dz = runtime.DeviceCells.from_host(host_arrays(cells), nx)
    st0 = gs.timestep(dz, g, synth.react_opts(dt=dt0), no, max_cuts=0)
    assert st0.reason == abi.PFX_SNES_DIVERGED_MAX_IT and st0.cuts == 0 and st0.dt == dt0
    gs.close()


def test_capped_derivative_gi_newton_converges_quadratically():
    # PFX_JAC_CAPPED_QK in the global Newton: the same converged state as
    # the reference derivative's Newton (the residual is the same), reached
    # in no more iterations
    t = synth.random_network(15, 50, 3, ionx=True)
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    rng = np.random.default_rng(61)
    q, dd = faces(rng, nx, ny, nz)
    cells, _, _, _ = speciated_state(t, n, seed=67)
    tab = runtime.Tables(t)
    g = runtime.DeviceGrid(nx, ny, nz, q, dd, log_formulation=True)
    gs = runtime.GiSolver(tab, g)
    # stol 1e-300: both stop on the residual (rtol 1e-12), not on a small
    # step, so both free-ion sets are converged
    no = abi.NewtonOpts(rtol=1e-12, stol=1e-300, max_iterations=80)
    no.krylov.rtol = 1e-14
    res = {}
    for jf in (0, abi.PFX_JAC_CAPPED_QK):
        dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
        st = gs.step(dc, g, synth.react_opts(jacobian_flags=jf), no)
        assert st.reason == abi.PFX_SNES_CONVERGED_FNORM_RELATIVE, (jf, st.as_dict())
        res[jf] = (st.iterations, dc.to_host()["pri_molal"])
    assert res[abi.PFX_JAC_CAPPED_QK][0] < res[0][0]
    assert np.max(np.abs(res[0][1] - res[1][1]) / res[0][1]) <= 1e-9
    gs.close()
