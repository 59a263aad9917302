"""Shared set-up of oracle-speciated cell states for the GPU parity tests."""
from oracle import speciated_state  # noqa: F401


def host_arrays(cells):
    return {k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}


def conditioning_floor(t, solved, step_inputs, dt, nthreads=0):
    """Per cell, the FP64 resolution of the free ions at a solution: the
    residual of a row cannot be evaluated to better than the unit roundoff
    times its largest term (the fixed accumulation over dt), and the Newton
    system in ln m maps that to eps * max_k sum_j |J_log^-1|_kj |fixed_j|/dt
    in ln m_k.  `solved` holds the solution (free ions and the step's
    activity coefficients), `step_inputs` the step's inputs (totals and
    sorbed totals of the previous state); opts as synth.react_opts."""
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
    s = np.abs(fixed) / dt
    with np.errstate(all="ignore"):
        for i in range(len(Jl)):
            try:
                Jinv = np.linalg.inv(Jl[i])
            except np.linalg.LinAlgError:
                continue   # singular: the free ions are not resolved at all
            fl[i] = 2.220446049250313e-16 * np.max(np.abs(Jinv) @ s[i])
    return np.where(np.isfinite(fl), fl, np.inf)
