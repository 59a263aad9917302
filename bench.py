"""Benchmark: batched per-cell reaction step on one MI355X.

Workload (the 1M-cell batched configuration): 100x100x100 = 1,000,000
independent cells, 15 primaries, 50 aqueous complexes (stoichiometry in
{-2..2}, log K ~ N(0,6), seeded van 't Hoff temperature dependence), 3
kinetic minerals, 1 ion-exchange site; totals log-uniform [1e-6, 1e-2]
mol/kg, pH U[5,9], T U[10,60] C; FP64; one reaction step of dt = 0.1 d per
cell.  Untimed set-up: the cells' previous state comes from the GPU
constraint speciation (pfx_equilibrate_batch) plus the initial aux-var
refresh.  Timed step: pfx_react_batch from the reference's default guess
(1e-9 molal, plain Newton) to rel_res <= 1e-10.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--ncells N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncells", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0, help="cells in the CPU baseline sample (0 = auto)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the per-cell reaction step shards by "
                                   "independent cell slabs; run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from pflotran_amd import abi, runtime, synth

    torch.cuda.set_device(0)
    t = synth.random_network(15, 50, 3, ionx=True)
    n = args.ncells
    tab = runtime.Tables(t)

    # ---- untimed set-up: previous state of every cell by GPU speciation;
    # cells whose speciation fails are replaced by further draws
    n_draw = int(n * 1.08) + 64
    ci = synth.random_cells(t, n_draw)
    arr = synth.host_state_arrays(t, n_draw, with_temperature=True)
    arr["temperature"][:] = ci.temp
    arr["mnrl_volfrac"][:] = ci.vf
    arr["mnrl_area"][:] = ci.area
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    ctype = [abi.PFX_CONSTRAINT_PH] + [abi.PFX_CONSTRAINT_TOTAL] * (t.N - 1)
    dc = runtime.DeviceCells.from_host(arr, n_draw)
    t0 = time.time()
    rc_eq, st_eq = runtime.equilibrate(tab, dc, ctype, conc, synth.react_opts(max_iterations=300))
    m = dc.t["pri_molal"]
    ok = (dc.t["status"] == 0) & torch.isfinite(m).all(dim=1) & (m > 0).all(dim=1)
    idx = torch.nonzero(ok).flatten()[:n]
    if idx.numel() < n:
        raise RuntimeError(f"only {idx.numel()} of {n} cells speciated")
    dc = runtime.DeviceCells({k: (None if v is None else v.index_select(0, idx).contiguous())
                              for k, v in dc.t.items()}, n)
    dc.t["pri_act_coef"].fill_(1.0)
    dc.t["sec_act_coef"].fill_(1.0)
    dc.t["ionx_ref_sorbed"].fill_(1e-9)
    runtime.auxvar(tab, dc, 2, synth.react_opts())
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"# set-up speciation: {n_draw} cells drawn, {n_draw - int(ok.sum())} failed, "
          f"{st_eq.kernel_ms:.1f} ms kernel, {setup_s:.2f} s", file=sys.stderr, flush=True)
    inputs = {k: dc.t[k].clone() for k in ("total", "total_sorb", "pri_molal", "pri_act_coef",
                                           "sec_act_coef", "ionx_ref_sorbed")}

    def restore():
        for k, v in inputs.items():
            dc.t[k].copy_(v)

    opts = synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT)
    stats = None
    for _ in range(args.warmup):
        restore()
        runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT))
    torch.cuda.synchronize()
    kern_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        restore()
        rc, st = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT))
        kern_ms.append(st.kernel_ms)
        stats = st
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    converged = stats.converged
    value = converged / (ms_per_step * 1e-3)

    # ---- roofline: algorithmic FP64 work of the Newton iterations
    fl = t.flops_per_iteration()
    iters = stats.total_iterations
    # the activity pass of each cell (one RTotal + Debye-Hueckel) counted as one iteration's RTotal
    flops = (iters + n) * fl["flops"] + (iters + n) * fl["transcendentals"]
    kern_avg_ms = float(np.mean(kern_ms))
    achieved_tflops = flops / (kern_avg_ms * 1e-3) / 1e12
    bytes_per_cell = 8 * (4 * t.N + 2 * t.M + 3 * t.K + t.R + 2) + 4 * 2
    line = {
        "metric": "reacted cells/s at 1e-10 rel residual: surrogate-warm vs plain GPU Newton",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based RNG, key 20261202)",
        "config": {"workload": f"random_network N=15 M=50 K=3 ionx=1, {n} cells, dt=8640 s, plain Newton "
                               f"from 1e-9 molal default guess (surrogate warm start not built yet)",
                   "ncells": n, "guess": "default 1e-9", "rel_tol": 1e-10},
        "newton": {"converged": int(converged), "total_iterations": int(iters),
                   "mean_iterations": iters / max(n, 1), "max_iterations": stats.max_iterations,
                   "min_iterations": stats.min_iterations, "max_rel_res": stats.max_rel_res,
                   "histogram": list(stats.iteration_histogram)},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "pfx_react_kernel<16,0>", "kernel_ms": kern_avg_ms,
                     "flops_per_iteration": fl["flops"], "transcendentals_per_iteration": fl["transcendentals"],
                     "note": "FP64 VALU-bound; MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s",
                     "algorithmic_bytes_per_cell": bytes_per_cell,
                     "achieved_GBs": bytes_per_cell * n / (kern_avg_ms * 1e-3) / 1e9},
    }
    if args.calibrate_copy:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
        # calibrate the byte counters (256 MiB read + 256 MiB written)
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = x.numel() * 8
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(t, args.cpu_sample)
    print(json.dumps(line), flush=True)


def cpu_baseline(t, sample):
    """The oracle (C restatement of RReact, OpenMP over cells) on a bounded
    sample of the same workload, timed on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    from pflotran_amd import synth

    threads = min(os.cpu_count() or 1, 16)
    n = sample or 12000 * threads
    cells, _, _, _ = orc.speciated_state(t, n, nthreads=threads)
    work = cells.copy()
    t0 = time.perf_counter()
    rc, st = orc.react_batch(t, work, synth.react_opts(), nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": st.converged / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"first {n} cells of the same seeded workload, same tables/tolerance/guess; "
                      f"{dt:.2f} s, {st.total_iterations} Newton iterations"}


if __name__ == "__main__":
    main()
