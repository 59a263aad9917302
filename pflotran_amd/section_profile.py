"""Where the Newton kernel's cycles go: per-section shader-clock totals from
the profiling build (libpfx_chem_prof.so, -DPFX_SECTION_TIMERS).

Workload: the 1M-cell batched network (15 primaries, 50 complexes, 3
minerals, 1 exchange site), one reaction step each of
  plain: from the default guess (1e-9 molal);
  warm:  from the converged free ions scaled by exp(N(0, 0.02)), a guess of
         the trained warm start's accuracy, with the fallback.
Each at the register cap of 3 and of 1 wavefronts per SIMD.  Cycles are
summed per wavefront over all its groups' passes, so a section's figure per
wavefront-iteration includes the time the SIMD spends on the other
wavefronts it holds (wall time of the wavefront, not issue time).

    python -m pflotran_amd.section_profile [ncells]

prints one JSON object.  Build first: python -m pflotran_amd.build --profiling
"""
from __future__ import annotations

import ctypes as C
import json
import sys

import numpy as np

from . import abi, build, runtime, synth, workload

SECTIONS = ["refill/load", "activity", "complexes", "totals+dT/dm", "ion exchange",
            "residual+minerals+convergence", "constraint rows", "LU solve", "update+retire"]
VARIANT_W = {0: 1, 2: 3}   # pfx_tables_tune variant -> wavefronts per SIMD


def main(n: int = 1_000_000):
    import torch
    lib = runtime.load_library(build.LIB_PROF)
    lib.pfx_section_cycles.argtypes = [C.c_void_p, C.c_int32]
    torch.cuda.set_device(0)
    t = synth.random_network(15, 50, 3, ionx=True)
    tab = runtime.Tables(t)
    dc, _ = workload.speciated_cells(t, tab, n)
    snap = workload.snapshot(dc)
    buf = (C.c_ulonglong * 16)()
    out = {"workload": f"{n} cells of the batched network N=15 M=50 K=3 ionx=1, one step (dt 0.1 d)",
           "command": "python -m pflotran_amd.section_profile", "runs": {}}

    def run(label, opts, variant, guess=None):
        tab.tune(abi.PFX_KERNEL_REACT, variant)
        workload.restore(dc, snap)
        if guess is not None:
            dc.t["guess"].copy_(guess)
        torch.cuda.synchronize()
        if lib.pfx_section_cycles(C.cast(buf, C.c_void_p), 1) != 0:
            raise runtime.PfxError("pfx_section_cycles failed")
        rc, st = runtime.react_batch(tab, dc, opts)
        torch.cuda.synchronize()
        if lib.pfx_section_cycles(C.cast(buf, C.c_void_p), 1) != 0:
            raise runtime.PfxError("pfx_section_cycles failed")
        cyc = np.array(buf[:9], dtype=np.float64)
        waves = int(buf[15])
        wave_its = st.total_iterations / 4.0      # 4 groups of 16 lanes per wavefront
        out["runs"][label] = {
            "waves_per_simd": VARIANT_W[variant], "kernel_ms": st.kernel_ms,
            "iterations_per_cell": st.total_iterations / n, "wavefronts": waves,
            "percent": {s: round(100.0 * c / cyc.sum(), 2) for s, c in zip(SECTIONS, cyc)},
            "cycles_per_wavefront_iteration": {s: round(c / wave_its) for s, c in zip(SECTIONS, cyc)}}
        return st

    plain = synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT)
    run("plain_w3", plain, 2)
    sol = dc.t["pri_molal"].clone()
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    guess = sol * torch.exp(0.02 * torch.randn(sol.shape, generator=g, device="cuda", dtype=torch.float64))
    warm = synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK, fallback_after=50)
    run("warm_w3", warm, 2, guess)
    run("plain_w1", plain, 0)
    run("warm_w1", warm, 0, guess)
    print(json.dumps(out, indent=1))
    return out


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1_000_000)
