"""Shared set-up of oracle-speciated cell states for the GPU parity tests."""
import numpy as np

import oracle as orc
from pflotran_amd import abi, synth


def speciated_state(t, n, seed=synth.SEED, workload_id=2, temperature=True, nthreads=0, only_ok=True):
    """Constraint speciation (totals of every primary but H+, pH) by the
    oracle: the previous state of each cell before its reaction step.  With
    only_ok, cells whose speciation fails are replaced by further draws."""
    if only_ok:
        cells, ctype, conc, ok = speciated_state(t, int(n * 1.2) + 8, seed, workload_id, temperature, nthreads,
                                                 only_ok=False)
        idx = np.where(ok)[0][:n]
        assert len(idx) == n, "not enough speciated cells"
        sub = orc.HostCells({k: (None if v is None else v[idx].copy()) for k, v in cells.arrays.items()}, n)
        return sub, ctype, conc[idx], np.ones(n, dtype=bool)
    ci = synth.random_cells(t, n, seed=seed, workload_id=workload_id)
    cells = orc.HostCells.allocate(t, n, temperature=temperature)
    if temperature:
        cells["temperature"][:] = ci.temp
    cells["mnrl_volfrac"][:] = ci.vf[:, :cells["mnrl_volfrac"].shape[1]]
    cells["mnrl_area"][:] = ci.area[:, :cells["mnrl_area"].shape[1]]
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    ctype = [orc.CON_PH] + [orc.CON_TOTAL] * (t.N - 1)
    o = synth.react_opts()
    rc, its = orc.equilibrate(t, ctype, conc, cells, o, nthreads=nthreads)
    ok = (its > 0) & np.all(np.isfinite(cells["pri_molal"]), axis=1) & np.all(cells["pri_molal"] > 0, axis=1)
    return cells, ctype, conc, ok


def host_arrays(cells):
    return {k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}
