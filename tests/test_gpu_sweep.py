"""Parity sweep of the reaction step over network sizes (N = 8, 15, 30;
with and without ion exchange) and both starting points (the reference's
default guess and the previous state): the GPU and the oracle must end
every cell in the same status, and co-converged cells' free ions must agree
to 1e-10.  Cells at the iteration cap stay at the cap in both."""
import numpy as np
import pytest

import oracle as orc
from pflotran_amd import abi, runtime, synth

pytestmark = pytest.mark.gpu

NETWORKS = [(15, 50, 3, True, 11), (8, 24, 3, False, 13), (30, 90, 3, False, 14)]


@pytest.mark.parametrize("N,M,K,ionx,seed", NETWORKS)
@pytest.mark.parametrize("mode", [abi.PFX_GUESS_DEFAULT, abi.PFX_GUESS_PREVIOUS])
def test_status_and_free_ions_match_the_oracle(N, M, K, ionx, seed, mode):
    t = synth.random_network(N, M, K, ionx=ionx, seed=synth.SEED + seed)
    drawn, _, _, ok = orc.speciated_state(t, 1500, seed=seed, only_ok=False)
    idx = np.where(ok)[0]
    n = len(idx)
    cells = orc.HostCells({k: (None if v is None else v[idx].copy()) for k, v in drawn.arrays.items()}, n)
    ref = cells.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=mode))
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host({k: (None if v is None else v.copy()) for k, v in cells.arrays.items()}, n)
    print(f"sweep N={N} M={M} ionx={ionx} mode={mode}: {n} cells", flush=True)
    runtime.react_batch(tab, dc, synth.react_opts(guess_mode=mode))
    out = dc.to_host()
    assert np.array_equal(out["status"], ref["status"])
    both = ref["status"] == 0
    assert both.sum() > 0.95 * n
    assert np.max(np.abs(out["pri_molal"][both] - ref["pri_molal"][both]) / ref["pri_molal"][both]) <= 1e-10
