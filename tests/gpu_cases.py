"""Shared set-up of oracle-speciated cell states for the GPU parity tests."""
from oracle import speciated_state  # noqa: F401


def host_arrays(cells):
    return {k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}


def row_term_scale(t, solved, fixed, o):
    """Per cell and row, the magnitude of the terms summed into the residual
    row (mol/s): the fixed accumulation, the free ion and every complex it
    enters (|nu| C) times phi s rho V, and the sorbed total times V, over
    dt; the denominator of the scaled residual."""
    import numpy as np
    absnu = np.zeros((t.M, t.N))
    for i in range(t.M):
        for s_ in range(t.cplx_specid[i, 0]):
            absnu[i, t.cplx_specid[i, 1 + s_]] += abs(t.cplx_stoich[i, s_])
    terms = solved["pri_molal"] + solved["sec_molal"][:, :t.M] @ absnu
    sc = o.porosity * o.saturation * o.den_kg * o.volume * terms + np.abs(fixed)
    if solved.arrays.get("total_sorb") is not None:
        sc = sc + np.abs(solved["total_sorb"]) * o.volume
    return sc / o.dt


def conditioning_floor(t, solved, step_inputs, dt, nthreads=0):
    """Per cell, the FP64 resolution of the free ions at a solution: the
    residual of a row cannot be evaluated to better than the unit roundoff
    times the magnitude of its terms (row_term_scale), and the Newton system
    in ln m maps that to eps * max_k sum_j |J_log^-1|_kj scale_j in ln m_k.
    `solved` holds the solution (free ions, the step's activity
    coefficients and the complexes), `step_inputs` the step's inputs (totals
    and sorbed totals of the previous state); opts as synth.react_opts."""
    import numpy as np

    import oracle as orc
    from pflotran_amd import synth
    o = synth.react_opts(dt=dt)
    fixed = o.porosity * o.saturation * 1000.0 * o.volume * step_inputs["total"]
    if step_inputs.arrays.get("total_sorb") is not None:
        fixed = fixed + step_inputs["total_sorb"] * o.volume
    ev = solved.copy()
    _, J = orc.residual_jacobian(t, ev, o, fixed_accum=np.ascontiguousarray(fixed), nthreads=nthreads)
    Jl = J * solved["pri_molal"][:, None, :]
    fl = np.full(len(Jl), np.inf)
    s = row_term_scale(t, ev, fixed, o)
    with np.errstate(all="ignore"):
        for i in range(len(Jl)):
            try:
                Jinv = np.linalg.inv(Jl[i])
            except np.linalg.LinAlgError:
                continue   # singular: the free ions are not resolved at all
            fl[i] = 2.220446049250313e-16 * np.max(np.abs(Jinv) @ s[i])
    return np.where(np.isfinite(fl), fl, np.inf)


def jacobian_bound(t, cells, jac_o):
    """Per-entry bound of the analytic Jacobian's rounding: 1e-12 of the
    magnitude of the terms summed into the entry (the complex terms
    |nu_r nu_k| C_c / m_k cancel in off-diagonal entries, so either
    implementation's rounding is relative to that sum, not to |J|); the
    exchanger rows additionally carry the inner Newton's own 1e-12 relative
    tolerance on KDj (reaction.F90:4407): 2e-12 of their row's largest entry.
    cells must hold sec_molal at the state."""
    import numpy as np

    from pflotran_amd import synth
    m = cells["pri_molal"]
    sec = cells["sec_molal"][:, :t.M]
    absnu = np.zeros((t.M, t.N))
    for i in range(t.M):
        for s_ in range(t.cplx_specid[i, 0]):
            absnu[i, t.cplx_specid[i, 1 + s_]] += abs(t.cplx_stoich[i, s_])
    psvd = 0.25 * 1.0 * 1000.0 * 1.0 / synth.STEP_DT * (synth.DEN_25C * 1e-3)
    terms = psvd * (np.eye(t.N)[None] + np.einsum("ci,ck,nc->nik", absnu, absnu, sec) / m[:, None, :])
    bound = 1e-12 * (terms + np.abs(jac_o))
    cation = np.zeros(t.N, dtype=bool)
    cation[t.ionx_cationid[0, 1:1 + t.ionx_cationid[0, 0]]] = True
    rowmax = np.max(np.abs(jac_o), axis=2, keepdims=True)
    bound[:, cation, :] = np.broadcast_to(2e-12 * rowmax, bound.shape)[:, cation, :]
    return bound
