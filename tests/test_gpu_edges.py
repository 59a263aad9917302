"""Edge cases of the batched entry points on the device: empty batches (no
cells, zero-size arrays), a full 32-lane group (N = 32 primaries: no padding
rows in the row-per-lane LU), and a network beyond the compiled limits."""
import numpy as np
import pytest

import oracle as orc
from gpu_cases import host_arrays, speciated_state
from pflotran_amd import abi, runtime, synth

pytestmark = pytest.mark.gpu


def test_empty_batches():
    t = synth.random_network(15, 50, 3, ionx=True)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(synth.host_state_arrays(t, 0, with_temperature=False), 0)
    print("empty batches", flush=True)
    rc, st = runtime.react_batch(tab, dc, synth.react_opts())
    assert rc == 0 and st.ncells == 0 and st.converged == 0 and st.total_iterations == 0
    rc, st = runtime.equilibrate(tab, dc, [abi.PFX_CONSTRAINT_TOTAL] * t.N, np.zeros((0, t.N)), synth.react_opts())
    assert rc == 0 and st.ncells == 0
    runtime.auxvar(tab, dc, 2, synth.react_opts())
    runtime.update_kinetic_state(tab, dc, synth.react_opts())


def test_full_group_of_32_primaries():
    """N = 32 fills the 32-lane group: every lane holds a live row, and the
    LU's steps all pivot live rows."""
    t = synth.random_network(32, 96, 3, ionx=False)
    drawn, _, _, ok = speciated_state(t, 192, seed=61, only_ok=False)
    idx = np.where(ok)[0]
    n = len(idx)
    assert n > 100  # about three in four draws of this network speciate
    cells = orc.HostCells({k: (None if v is None else v[idx].copy()) for k, v in drawn.arrays.items()}, n)
    ref = cells.copy()
    rc_o, _ = orc.react_batch(t, ref, synth.react_opts())
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    print("N = 32 react", flush=True)
    rc_g, st = runtime.react_batch(tab, dc, synth.react_opts())
    out = dc.to_host()
    assert np.array_equal(out["status"], ref["status"])
    conv = ref["status"] == 0
    assert conv.sum() > 0.9 * n
    assert np.max(np.abs(out["pri_molal"][conv] - ref["pri_molal"][conv]) / ref["pri_molal"][conv]) <= 1e-10


def test_network_beyond_the_limits_is_refused():
    t = synth.random_network(33, 40, 1, ionx=False)
    with pytest.raises(runtime.PfxError):
        runtime.Tables(t)
