"""pfx_react_batch_noalloc (the entry point that allocates nothing: its
queue counters, statistics and per-cell scratch live in the library's fixed
device array) against pfx_react_batch on the same inputs: identical status,
free ions, totals, coefficients, rates and statistics, on two calls in a
row, the second with fewer cells (the array keeps the first call's
contents), and the size limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from gpu_cases import host_arrays, speciated_state
from pflotran_amd import abi, runtime, synth

pytestmark = pytest.mark.gpu

OUT = ("status", "num_iterations", "pri_molal", "total", "total_sorb", "pri_act_coef", "sec_act_coef", "sec_molal",
       "mnrl_rate", "ionx_ref_sorbed", "rel_res")


def _stats_tuple(st):
    return (st.ncells, st.converged, st.total_iterations, st.max_iterations, st.min_iterations, st.max_rel_res,
            tuple(st.iteration_histogram), st.fallbacks)


@pytest.mark.parametrize("per_cell_outputs", [False, True])
def test_noalloc_matches_react_batch_on_two_calls(per_cell_outputs):
    t = synth.random_network(15, 50, 3, ionx=True)
    tab = runtime.Tables(t)
    for n, seed in ((3000, 71), (777, 73)):
        cells, _, _, _ = speciated_state(t, n, seed=seed)
        opts = dict(guess_mode=abi.PFX_GUESS_PREVIOUS)
        ref = runtime.DeviceCells.from_host(host_arrays(cells), n)
        rc_r, st_r = runtime.react_batch(tab, ref, synth.react_opts(**opts))
        arr = host_arrays(cells)
        if not per_cell_outputs:
            for k in ("num_iterations", "status", "rel_res"):
                arr[k] = None
        dc = runtime.DeviceCells.from_host(arr, n)
        rc_n, st_n = runtime.react_batch_noalloc(tab, dc, synth.react_opts(**opts))
        assert rc_n == rc_r >= 0
        assert _stats_tuple(st_n) == _stats_tuple(st_r)
        a, b = dc.to_host(), ref.to_host()
        for k in OUT:
            if a.get(k) is None:
                continue
            assert np.array_equal(a[k], b[k]), k


def test_noalloc_refuses_batches_beyond_its_scratch():
    # no per-cell outputs and more cells than the fixed array holds: the
    # call returns PFX_E_ARG before it launches anything
    t = synth.random_network(8, 12, 0, ionx=False)
    tab = runtime.Tables(t)
    cells, _, _, _ = speciated_state(t, 4, seed=5)
    arr = host_arrays(cells)
    for k in ("num_iterations", "status", "rel_res"):
        arr[k] = None
    dc = runtime.DeviceCells.from_host(arr, 4)
    cs = dc.struct()
    cs.ncells = abi.PFX_NOALLOC_MAX_CELLS + 1
    lib = runtime.load_library()
    st = abi.Stats()
    o = synth.react_opts()
    assert lib.pfx_react_batch_noalloc(tab.handle, C.byref(cs), C.byref(o), C.byref(st)) == abi.PFX_E_ARG
    # a stream-ordered request is refused too (the statistics are read back)
    cs.ncells = 4
    o2 = synth.react_opts(stream_ordered=1)
    assert lib.pfx_react_batch_noalloc(tab.handle, C.byref(cs), C.byref(o2), C.byref(st)) == abi.PFX_E_ARG
