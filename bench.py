"""Benchmark: batched per-cell reaction step on one MI355X, learned warm start
against the plain Newton.

Workload (the 1M-cell batched configuration): 100x100x100 = 1,000,000
independent cells, 15 primaries, 50 aqueous complexes (stoichiometry in
{-2..2}, log K ~ N(0,6), seeded van 't Hoff temperature dependence), 3
kinetic minerals, 1 ion-exchange site; totals log-uniform [1e-6, 1e-2]
mol/kg, pH U[5,9], T U[10,60] C; FP64; one reaction step of dt = 0.1 d per
cell, converged to rel_res <= 1e-10.

Untimed set-up: the held-out cells' previous state (GPU constraint
speciation + initial aux-var refresh); the warm-start network
40-256-256-256-15 (GELU, BF16 matrix cores, FP32 accumulation) is trained on
the GPU on pairs that the GPU solver generates from training cells drawn with
a different RNG key.

Timed, per step (inputs already in HBM):
  plain_previous: pfx_react_batch from the previous free ions (RReact's own
         starting point, reaction.F90:3370): the plain GPU Newton;
  plain: pfx_react_batch from the constraint default guess (1e-9 molal,
         reaction.F90:1481);
  warm:  pfx_surrogate_guess (features + network + guess, one launch), then
         pfx_react_batch from that guess, a failed attempt repeating the
         step from the previous free ions.
All three stop at rel_res <= 1e-10 with the step test (Newton step at the
accepted iterate below 1e-11 relative).  `value` is the warm method's
converged cells/s; both plain figures sit beside it.  The CPU baseline is
the oracle (OpenMP C restatement of RReact) from the previous free ions on
the first cells of the same held-out batch, timed on every host core of the
GPU's share and on one core, and its free ions are compared with the GPU's.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--ncells N] [--no-cpu]
                [--train-pairs P] [--epochs E] [--no-surrogate]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def timed_steps(torch, k, body):
    """k steps of body(); per-step wall ms (device synchronised each step)."""
    ms = []
    torch.cuda.synchronize()
    for _ in range(k):
        t0 = time.perf_counter()
        body()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    return ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncells", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0, help="cells in the CPU baseline sample (0 = auto)")
    ap.add_argument("--train-pairs", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=float, default=10.0)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--fallback-after", type=int, default=100,
                    help="warm attempt budget before the restart from the previous free ions")
    ap.add_argument("--no-surrogate", action="store_true", help="plain method only")
    ap.add_argument("--no-gi", action="store_true", help="skip the global-implicit assembly leg")
    ap.add_argument("--no-zones", action="store_true", help="skip the chemistry-zones leg")
    ap.add_argument("--no-column", action="store_true", help="skip the 100-cell operator-split column leg")
    ap.add_argument("--no-capped", action="store_true",
                    help="skip the legs with the consistent derivative above the Q/K cap")
    ap.add_argument("--zones-scale", type=float, default=1.0, help="fraction of the 2M zoned cells")
    ap.add_argument("--gi-side", type=int, default=128, help="cells per side of the global-implicit grid")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the per-cell reaction step shards by "
                                   "independent cell slabs; run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from pflotran_amd import abi, autotune, runtime, surrogate, synth, workload

    torch.cuda.set_device(0)
    t = synth.random_network(15, 50, 3, ionx=True)
    n = args.ncells
    tab = runtime.Tables(t)

    # ---- untimed set-up: held-out cells (RNG key 2)
    t0 = time.time()
    dc, nfail = workload.speciated_cells(t, tab, n)
    torch.cuda.synchronize()
    log(f"held-out set-up: {n} cells, {nfail} failed speciation draws replaced, {time.time() - t0:.2f} s")
    snap = workload.snapshot(dc)

    def restore():
        workload.restore(dc, snap)

    box = {}
    n_cpu = min(args.cpu_sample or 12000 * cpu_threads(), n)
    gpu_out = {}   # free ions and status of each GPU method on the CPU baseline's cells

    def plain_leg(key, opts):
        """one plain method: autotuned kernel variant, warm-up, timed steps"""
        def step():
            restore()
            box["rc"], box["st"] = runtime.react_batch(tab, dc, opts)
        # the trials start from the step's inputs (not from the state a
        # previous leg left)
        restore()
        tune = autotune.tune_kernel(tab, dc, lambda: runtime.react_batch(tab, dc, opts))
        log(f"{key} kernel variants (ms): {tune}")
        for _ in range(args.warmup):
            step()
        ms = timed_steps(torch, args.steps, step)
        st = box["st"]
        gpu_out[key] = (dc.t["pri_molal"][:n_cpu].cpu().numpy().copy(), dc.t["status"][:n_cpu].cpu().numpy().copy())
        ms_step = float(np.mean(ms))
        value = st.converged / (ms_step * 1e-3)
        log(f"{key}: {ms_step:.2f} ms/step, {value:.4g} cells/s, {st.total_iterations / n:.2f} it/cell")
        return {"cells_per_s": value, "ms_per_step": ms_step, "kernel_ms": st.kernel_ms,
                "converged": int(st.converged), "mean_iterations": st.total_iterations / n,
                "max_iterations": st.max_iterations, "max_rel_res": st.max_rel_res,
                "histogram": list(st.iteration_histogram), "kernel_variant_ms": tune,
                "kernel_variant": tab.tune(abi.PFX_KERNEL_REACT)}, st

    # ---- plain GPU Newton from the previous free ions (RReact's start) and
    # from the constraint default guess
    plain_prev, st_prev = plain_leg("plain_previous", synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
    plain, st_plain = plain_leg("plain_default", synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT))
    plain_value = plain["cells_per_s"]

    line = {
        "metric": "reacted cells/s at 1e-10 rel residual: surrogate-warm vs plain GPU Newton",
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based RNG, seed 20261202; held-out cells key 2, training cells key 4)",
    }

    fl = t.flops_per_iteration()
    warm = None
    mlp = None
    if not args.no_surrogate:
        # ---- untimed: training pairs from the GPU solver, then training
        mlp = surrogate.MLP(surrogate.DIMS, batch_max=args.batch, seed=synth.SEED)
        t0 = time.time()
        X_raw, Y, gst = surrogate.generate_pairs(t, tab, mlp, args.train_pairs, chunk=min(1_000_000, n), log=log)
        Xn, Yn = surrogate.fit_normalisation(mlp, X_raw, Y)
        del X_raw
        log(f"training pairs: {Xn.shape[0]} in {time.time() - t0:.1f} s")
        # hold 1% of the pairs out of training for the prediction error
        n_val = min(100_000, Xn.shape[0] // 100)
        tr = surrogate.train(mlp, Xn[n_val:], Yn[n_val:], epochs=args.epochs, batch=args.batch, lr=args.lr,
                             log=log, log_every=1000)
        pred = surrogate.prediction_error(mlp, Xn[:n_val], Y[:n_val])
        del Xn, Yn, Y
        torch.cuda.empty_cache()
        log(f"trained: {tr['steps']} steps in {tr['seconds']:.1f} s, prediction rms {pred['rms_ln']:.3f} ln units")

        warm_opts = synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK, fallback_after=args.fallback_after,
                                     fallback_previous=1)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        mlp_ms = []

        def warm_step():
            restore()
            ev[0].record()
            surrogate.guess(tab, mlp, dc)
            ev[1].record()
            box["rc"], box["st"] = runtime.react_batch(tab, dc, warm_opts)
            mlp_ms.append(ev[0].elapsed_time(ev[1]))

        restore()
        surrogate.guess(tab, mlp, dc)
        warm_tune = autotune.tune_kernel(tab, dc, lambda: runtime.react_batch(tab, dc, warm_opts))
        log(f"warm kernel variants (ms): {warm_tune}")
        for _ in range(args.warmup):
            warm_step()
        mlp_ms.clear()
        warm_ms = timed_steps(torch, args.steps, warm_step)
        st_warm = box["st"]
        warm_ms_step = float(np.mean(warm_ms))
        warm_value = st_warm.converged / (warm_ms_step * 1e-3)
        log(f"warm: {warm_ms_step:.2f} ms/step, {warm_value:.4g} cells/s, "
            f"{st_warm.total_iterations / n:.2f} it/cell, {st_warm.fallbacks} fallbacks")
        mlp_avg = float(np.mean(mlp_ms))
        mlp_flops = 2 * sum(a * b for a, b in zip(surrogate.DIMS[:-1], surrogate.DIMS[1:]))
        warm = {"cells_per_s": warm_value, "ms_per_step": warm_ms_step, "kernel_ms": st_warm.kernel_ms,
                "surrogate_ms": mlp_avg, "converged": int(st_warm.converged),
                "mean_iterations": st_warm.total_iterations / n, "max_iterations": st_warm.max_iterations,
                "max_rel_res": st_warm.max_rel_res, "fallbacks": int(st_warm.fallbacks),
                "fallback_rate": st_warm.fallbacks / n, "fallback_after": args.fallback_after,
                "histogram": list(st_warm.iteration_histogram),
                "speedup_vs_plain_previous": warm_value / plain_prev["cells_per_s"],
                "iteration_reduction_vs_plain_previous": st_prev.total_iterations / max(st_warm.total_iterations, 1),
                "speedup_vs_plain_default": warm_value / plain_value,
                "iteration_reduction_vs_plain_default": st_plain.total_iterations / max(st_warm.total_iterations, 1),
                "kernel_variant_ms": warm_tune, "kernel_variant": tab.tune(abi.PFX_KERNEL_REACT),
                "prediction_error_ln": pred,
                "training": {"pairs": int(args.train_pairs), "epochs": args.epochs, "batch": args.batch,
                             "steps": tr["steps"], "seconds": tr["seconds"], "samples_per_s": tr["samples_per_s"],
                             "final_loss": tr["loss"][-1][1], "pair_generation_s": gst["seconds"]},
                "mlp_roofline": {"bound": "mfma", "flops_per_cell": mlp_flops,
                                 "achieved": mlp_flops * n / (mlp_avg * 1e-3) / 1e12, "peak": PEAK_BF16_TFLOPS,
                                 "unit": "TFLOP/s",
                                 "frac": mlp_flops * n / (mlp_avg * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
                                 "kernel": "k_surrogate_guess"}}
        value, ms_per_step, st_main, kern_ms = warm_value, warm_ms_step, st_warm, st_warm.kernel_ms
        gpu_m = dc.t["pri_molal"].clone()
        gpu_status = dc.t["status"].clone()
        gpu_out["warm"] = (gpu_m[:n_cpu].cpu().numpy().copy(), gpu_status[:n_cpu].cpu().numpy().copy())

        # both methods again with PFX_JAC_CAPPED_QK (no mineral derivative
        # above the Q/K cap, where the rate is constant): same residual and
        # acceptance test, a consistent Jacobian for the capped cells
        cap = {}
        for key, opts in () if args.no_capped else (("plain_previous", synth.react_opts(
                                                         guess_mode=abi.PFX_GUESS_PREVIOUS,
                                                         jacobian_flags=abi.PFX_JAC_CAPPED_QK)),
                          ("warm", synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK,
                                                    fallback_after=args.fallback_after, fallback_previous=1,
                                                    jacobian_flags=abi.PFX_JAC_CAPPED_QK))):
            def step(opts=opts, key=key):
                restore()
                if key == "warm":
                    surrogate.guess(tab, mlp, dc)
                box["rc"], box["st"] = runtime.react_batch(tab, dc, opts)
            for _ in range(args.warmup):
                step()
            ms = float(np.mean(timed_steps(torch, args.steps, step)))
            st = box["st"]
            cap[key] = {"cells_per_s": st.converged / (ms * 1e-3), "ms_per_step": ms, "kernel_ms": st.kernel_ms,
                        "converged": int(st.converged), "mean_iterations": st.total_iterations / n,
                        "max_iterations": st.max_iterations, "max_rel_res": st.max_rel_res,
                        "fallbacks": int(st.fallbacks)}
            log(f"{key} with PFX_JAC_CAPPED_QK: {ms:.2f} ms/step, {cap[key]['cells_per_s']:.4g} cells/s, "
                f"{cap[key]['mean_iterations']:.2f} it/cell")
        if cap:
            cap["warm_speedup_vs_plain_previous"] = cap["warm"]["cells_per_s"] / cap["plain_previous"]["cells_per_s"]
            cap["warm_speedup_vs_reference_jacobian_plain_previous"] = (cap["warm"]["cells_per_s"]
                                                                        / plain_prev["cells_per_s"])
            both = ((dc.t["status"] == 0) & (gpu_status == 0)).unsqueeze(1)
            rel = ((dc.t["pri_molal"] - gpu_m).abs() / gpu_m.abs()).masked_fill(~both, 0.0)
            cap["max_rel_diff_free_ions_vs_reference_jacobian_warm"] = float(rel.max().item())
            cap["p99_rel_diff_free_ions_vs_reference_jacobian_warm"] = float(
                torch.quantile(rel.amax(dim=1)[:200_000].float(), 0.99).item())
            warm["capped_qk"] = cap
    else:
        value, ms_per_step, st_main, kern_ms = (plain_prev["cells_per_s"], plain_prev["ms_per_step"], st_prev,
                                                plain_prev["kernel_ms"])

    # ---- roofline of the dominant kernel (the FP64 Newton): algorithmic
    # flops of the Newton updates plus, per cell, the final evaluation and
    # the LU solve of its step test (the RTotal + activity pass at the
    # previous state is not counted)
    iters = st_main.total_iterations
    flops = (iters + n) * (fl["flops"] + fl["transcendentals"])
    achieved = flops / (kern_ms * 1e-3) / 1e12
    # algorithmic bytes per cell: read totals, sorbed totals, free ions,
    # guess, both activity-coefficient arrays, vf, area, T and the exchange
    # state once; write free ions, totals, sorbed totals, coefficients,
    # complexes, rates, exchange state, rel_res, iterations and status once
    bytes_per_cell = 8 * ((5 * t.N + t.M + 2 * t.K + 1 + t.R) + (4 * t.N + 2 * t.M + t.K + t.R + 1)) + 8
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic_pfx_react_kernel_1M.json")) as f:
            prof = json.load(f)["pfx_react_kernel<16,0>"]
        # the counters were read on the uncapped variant: they stand for
        # the timed kernel only when the autotuner picked that one too
        if n == 1_000_000 and warm is not None and tab.tune(abi.PFX_KERNEL_REACT) == 0:
            traffic = prof["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        traffic = None
    line.update({
        "value": value,
        "ms_per_step": ms_per_step,
        "config": {"workload": f"random_network N=15 M=50 K=3 ionx=1, {n} held-out cells, dt=8640 s, rel_tol 1e-10; "
                               + ("warm start 40-256-256-256-15 BF16 MLP + FP64 Newton with fallback"
                                  if warm else "plain Newton from the previous free ions"),
                   "ncells": n, "rel_tol": 1e-10, "rel_change_tol": synth.REL_CHANGE_TOL,
                   "timed_step": ("restore inputs + pfx_surrogate_guess + pfx_react_batch(ARRAY_FALLBACK, "
                                  "fallback_previous)" if warm else "restore inputs + pfx_react_batch(PREVIOUS)")},
        "plain_previous": plain_prev,
        "plain": plain,
        "warm": warm,
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "traffic_source": "profiles/pmc_traffic_pfx_react_kernel_1M.json (FETCH_SIZE x2.0, "
                                       "WRITE_SIZE x1.0 from a 256 MiB copy)" if traffic else None,
                     "kernel": f"pfx_react_kernel<16, 0, {tab.tune(abi.PFX_KERNEL_REACT) + 1}, 0, false> "
                               "(unpivoted) + pfx_react_kernel<16, 0, 1, 1, false> (pivoting follow-up)",
                     "kernel_ms": kern_ms,
                     "flops_per_iteration": fl["flops"], "transcendentals_per_iteration": fl["transcendentals"],
                     "note": "FP64 VALU-bound Newton; MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s",
                     "algorithmic_bytes_per_cell": bytes_per_cell,
                     "achieved_GBs": bytes_per_cell * n / (kern_ms * 1e-3) / 1e9},
    })
    if not args.no_gi:
        line["gi"] = gi_leg(t, tab, args, log, mlp=mlp)
    if not args.no_zones:
        line["zones"] = zones_leg(args, log)
    if not args.no_column:
        line["column_os"] = column_leg(args, log)
    if args.calibrate_copy:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
        # calibrate the byte counters (256 MiB read + 256 MiB written)
        x = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = x.numel() * 8
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        line["copy_GBs"] = 2 * x.numel() * 8 * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(t, dc, snap, gpu_out, n_cpu)
    print(json.dumps(line), flush=True)


def gi_leg(t, tab, args, log, mlp=None):
    """Global-implicit assembly of a side^3 grid (7-point, upwind): the
    first residual/Jacobian evaluation of a time step (x = old state), the
    BSR Jacobian with column-major 15x15 blocks.  HBM-bound."""
    import numpy as np
    import torch

    from pflotran_amd import abi, runtime, synth, workload

    side = args.gi_side
    n = side ** 3
    torch.cuda.empty_cache()
    t0 = time.time()
    dc, _ = workload.speciated_cells(t, tab, n, workload_id=5)
    bc_total = dc.t["total"].mean(dim=0).cpu().numpy()
    g = runtime.DeviceGrid(**workload.gi_grid(side, side, side, bc_total=bc_total))
    row_ptr, col_idx = runtime.gi_structure(g)
    nnzb = int(col_idx.numel())
    N = t.N
    res = torch.empty((n, N), dtype=torch.float64, device="cuda")
    vals = torch.empty((nnzb, N, N), dtype=torch.float64, device="cuda")
    log(f"gi set-up: {n} cells, {nnzb} blocks ({vals.numel() * 8 / 1e9:.1f} GB), {time.time() - t0:.1f} s")
    opts = synth.react_opts()
    runtime.gi_residual_jacobian(tab, dc, g, opts, row_ptr, res, vals)   # warm-up
    ev, asm, wall = [], [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        runtime.gi_residual_jacobian(tab, dc, g, synth.react_opts(), row_ptr, res, vals)
        torch.cuda.synchronize()
        wall.append((time.perf_counter() - t1) * 1e3)
        e_ms, a_ms = tab.timing()
        ev.append(e_ms)
        asm.append(a_ms)
    a_ms = float(np.mean(asm))
    # algorithmic bytes of the assembly kernel: per cell read the non-flux
    # residual, totals, non-flux block and dT/dm once; write the residual and
    # every BSR block once
    bytes_asm = 8 * (n * (2 * N + 2 * N * N) + n * N + nnzb * N * N)
    ok = bool(torch.isfinite(res).all()) and bool(torch.isfinite(vals).all())
    out = {"workload": f"{side}^3 grid, N={N}, 7-point upwind, ln k ~ N(ln 1e-12, 1), log unknowns",
           "ncells": n, "nnzb": nnzb, "ms_per_evaluation": float(np.mean(wall)),
           "ms_cell_kernel": float(np.mean(ev)), "ms_assembly_kernel": a_ms,
           "assembly_bytes": bytes_asm, "assembly_GBs": bytes_asm / (a_ms * 1e-3) / 1e9,
           "assembly_frac_hbm": bytes_asm / (a_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "cells_per_s": n / (float(np.mean(wall)) * 1e-3), "finite": ok}
    log(f"gi: {out['ms_per_evaluation']:.2f} ms/evaluation, assembly {a_ms:.2f} ms "
        f"({out['assembly_GBs']:.0f} GB/s)")

    # the block product of the Krylov solve on this Jacobian: HBM-bound,
    # algorithmic bytes = every block once, x and y once, the structure once
    bsr = runtime.Bsr(row_ptr, col_idx, N)
    xv = torch.rand((n, N), dtype=torch.float64, device="cuda")
    yv = torch.empty_like(xv)
    bsr.spmv(vals, xv, yv)
    sp_ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        bsr.spmv(vals, xv, yv)
        e1.record()
        torch.cuda.synchronize()
        sp_ms.append(e0.elapsed_time(e1))
    spmv_bytes = 8 * (nnzb * N * N + 2 * n * N) + 4 * (nnzb + n + 1)
    out["spmv_ms"] = float(np.mean(sp_ms))
    out["spmv_GBs"] = spmv_bytes / (out["spmv_ms"] * 1e-3) / 1e9
    out["spmv_frac_hbm"] = out["spmv_GBs"] / PEAK_HBM_GBS
    bsr.close()
    del vals, res, xv, yv
    torch.cuda.empty_cache()
    log(f"gi: block product {out['spmv_ms']:.2f} ms ({out['spmv_GBs']:.0f} GB/s)")

    # one time step closed on the device: Newton from the old state with
    # point-block Jacobi BiCGStab, PETSc's default tolerances, and the
    # reference's time-step cut (halve dt and retry from the old state, at
    # most 16 times) when Newton does not converge; with the reference's
    # reaction Jacobian, with the capped-Q/K derivative (first 12 Newton
    # iterations), and with NUMERICAL_JACOBIAN
    # The same step from a predicted iterate (runtime.gi_initial_iterate):
    # every cell's own reaction step first (pfx_react_batch, from the old
    # free ions or from the learned warm start), then the global Newton from
    # its free ions; same residual, Jacobian and tests, so the same solution
    # within SNES's tolerances.  Timed: predictor + global step.
    snap = workload.snapshot(dc)
    ref_m = {}
    cases = (("newton", 0, 0, None), ("newton_capped_qk", 0, abi.PFX_JAC_CAPPED_QK, None),
             ("newton_numerical_jacobian", 1, 0, None), ("newton_predicted_previous", 0, 0, "previous"),
             ("newton_predicted_warm", 0, 0, "warm"),
             ("newton_capped_qk_predicted_warm", 0, abi.PFX_JAC_CAPPED_QK, "warm"))
    for key, numerical, jflags, pred in cases:
        if pred == "warm" and mlp is None:
            continue
        workload.restore(dc, snap)
        t.numerical_jacobian = numerical
        tab_n = runtime.Tables(t)
        t.numerical_jacobian = 0
        gs = runtime.GiSolver(tab_n, g)
        dt_want = synth.STEP_DT
        o = synth.react_opts(dt=dt_want, jacobian_flags=jflags)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pst = None
        if pred is not None:
            warm = None
            if pred == "warm":
                from pflotran_amd import surrogate

                def warm(tb, cells):
                    surrogate.guess(tb, mlp, cells)
            pst = runtime.gi_initial_iterate(tab_n, dc, o, warm=warm)
            torch.cuda.synchronize()
        pred_ms = (time.perf_counter() - t1) * 1e3
        st = gs.timestep(dc, g, o)
        torch.cuda.synchronize()
        total_ms = (time.perf_counter() - t1) * 1e3
        gs.close()
        ns = st.newton
        out[key] = {"reason": st.reason, "converged": st.reason > 0, "dt_wanted": dt_want, "dt_taken": st.dt,
                    "cuts": st.cuts, "newton_iterations": st.newton_iterations,
                    "wasted_newton_iterations": st.wasted_newton_iterations,
                    "linear_iterations": st.linear_iterations, "ms_total": total_ms,
                    "last_solve": {"iterations": ns.iterations, "fnorm0": ns.fnorm0, "fnorm": ns.fnorm,
                                   "ms_assembly": ns.ms_assembly, "ms_linear": ns.ms_linear},
                    "ms_per_newton_iteration": st.ms_total / max(st.newton_iterations, 1),
                    "tolerances": "SNES rtol 1e-8 / stol 1e-8 / 50 its, KSP rtol 1e-5 (PETSc defaults); "
                                  "cut factor 0.5, at most 16 cuts (timestepper defaults)"}
        if pst is not None:
            out[key]["predictor"] = {"start": pred, "ms": pred_ms, "converged_cells": int(pst.converged),
                                     "mean_iterations": pst.total_iterations / n}
        if st.reason <= 0:
            out[key]["note"] = "NOT CONVERGED after the cut budget: the time is not a step time"
        m_end = dc.t["pri_molal"].clone()
        jkey = "capped" if jflags else "reference"
        if pred is None and not numerical and st.dt == dt_want and st.reason > 0:
            ref_m[jkey] = m_end
        if pred is not None and jkey in ref_m:
            rel = ((m_end - ref_m[jkey]).abs() / ref_m[jkey].abs()).amax(dim=1)
            out[key]["free_ions_vs_unpredicted_step"] = {
                "max_rel": float(rel.max().item()),
                "p99_rel": float(torch.quantile(rel[:1_000_000].float(), 0.99).item())}
        del m_end
        log(f"gi: {key} {total_ms:.1f} ms (predictor {pred_ms:.1f}), reason {st.reason}, dt {st.dt:.0f} s after "
            f"{st.cuts} cuts, {st.newton_iterations} Newton its ({st.wasted_newton_iterations} wasted), "
            f"{st.linear_iterations} BiCGStab, last |F| {ns.fnorm0:.2e} -> {ns.fnorm:.2e}")
        del tab_n
    del ref_m

    # the transport half of an operator-split step on the same grid: the
    # scalar transport matrix once, one BiCGStab solve per component (Jacobi,
    # KSP rtol 1e-8), dt 1 day; the matrix product at block size 1 reads
    # 12 B per non-zero plus x and y
    workload.restore(dc, snap)
    tr = runtime.OsTransport(g)
    o = abi.default_opts(dt=86400.0)
    tr.step(dc, g, N, o, rtol=1e-8)   # warm-up
    workload.restore(dc, snap)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rc, ks = tr.step(dc, g, N, o, rtol=1e-8)
    torch.cuda.synchronize()
    os_ms = (time.perf_counter() - t1) * 1e3
    tr.close()
    out["os_transport"] = {"ms_step": os_ms, "components": N, "krylov_iterations": ks.iterations,
                           "reason": ks.reason, "ms_krylov_device": ks.ms,
                           "cell_components_per_s": n * N / (os_ms * 1e-3),
                           "spmv_bytes_per_iteration": 12 * nnzb + 16 * n,
                           "note": "implicit transport of every component's totals (pfx_os_transport_step), "
                                   "dt 1 d, Jacobi BiCGStab rtol 1e-8"}
    log(f"gi: operator-split transport step {os_ms:.1f} ms for {N} components, {ks.iterations} BiCGStab its, "
        f"reason {ks.reason}")
    del dc, snap
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_gi(t, tab, log)
    return out


def cpu_baseline_gi(t, tab, log, side=16):
    """The global-implicit assembly against the oracle on a side^3 box
    drawn from the same distributions as the 128^3 grid (cells, log-normal
    permeability, boundary): relative errors of the residual and of the BSR
    blocks, and the oracle's assembly rate on every host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import torch

    import oracle as orc
    from pflotran_amd import runtime, synth, workload
    n = side ** 3
    dc, _ = workload.speciated_cells(t, tab, n, workload_id=5)
    bc_total = dc.t["total"].mean(dim=0).cpu().numpy()
    gd = workload.gi_grid(side, side, side, bc_total=bc_total)
    g = runtime.DeviceGrid(**gd)
    rp, ci = runtime.gi_structure(g)
    N = t.N
    res = torch.empty((n, N), dtype=torch.float64, device="cuda")
    vals = torch.empty((int(ci.numel()), N, N), dtype=torch.float64, device="cuda")
    runtime.gi_residual_jacobian(tab, dc, g, synth.react_opts(), rp, res, vals)
    host = workload.host_arrays(dc)
    cells = orc.HostCells(host, n)
    hg = orc.HostGrid(gd["nx"], gd["ny"], gd["nz"], gd["q"], gd["dd"], area=gd["area"], bc_q=gd["bc_q"],
                      bc_dd=gd["bc_dd"], bc_total=gd["bc_total"], log_formulation=gd["log_formulation"])
    threads = cpu_threads()
    t0 = time.perf_counter()
    r_o, v_o, _, _ = orc.gi_residual_jacobian(t, cells, hg, synth.react_opts(), nthreads=threads)
    cpu_s = time.perf_counter() - t0
    r_g = res.cpu().numpy()
    v_g = np.ascontiguousarray(vals.cpu().numpy().transpose(0, 2, 1))   # [block][row][col] as the oracle's
    acc = np.abs(r_o) + 0.25 * 1000.0 * np.abs(host["total"]) / synth.STEP_DT
    rowmax = np.max(np.abs(v_o), axis=2, keepdims=True)
    out = {"box": f"{side}^3 box drawn from the same distributions", "cores": threads, "kind": "port",
           "oracle_assembly_cells_per_s": n / cpu_s,
           "residual_max_rel_err": float(np.max(np.abs(r_g - r_o) / acc)),
           "jacobian_max_err_rel_to_row_max": float(np.max(np.abs(v_g - v_o) / np.maximum(rowmax, 1e-300)))}
    log(f"gi vs oracle on {side}^3: {out}")
    return out


def column_leg(args, log):
    """The CPU-runnable configuration: a 1-D column of 100 cells (dx 1 m,
    porosity 0.25, Darcy flux 1 m/yr), 100 operator-split steps of 0.1 d,
    carbonate chemistry (5 primaries, 8 complexes, kinetic calcite from the
    hanford database excerpt), seeded totals log-uniform [1e-5, 1e-2] and pH
    U[6, 9] speciated on the device.  Each step: the explicit upwind
    transport of the totals, the batched reaction step from the previous free
    ions, the volume-fraction update.  Timed on the device (the whole run,
    launch-bound at 100 cells); the oracle's run of the same start on one
    core is the CPU baseline beside it."""
    import numpy as np
    import torch

    from pflotran_amd import abi, column, runtime, synth
    from pflotran_amd.database import MineralKinetics, build_tables, read_database

    db = read_database(os.path.join(ROOT, "tests", "golden", "db", "hanford_subset.dat"))
    t = build_tables(db, ["H+", "HCO3-", "Ca++", "Mg++", "Cl-"],
                     ["OH-", "CO2(aq)", "CO3--", "CaHCO3+", "CaCO3(aq)", "MgHCO3+", "MgCO3(aq)", "CaCl+"],
                     {"Calcite": MineralKinetics(1e-9)}, act_freq=abi.PFX_ACT_TIMESTEP)
    col = column.Column()
    n, nsteps, dt = col.ncells, 100, 8640.0
    tab = runtime.Tables(t)
    ci = synth.random_cells(t, n + 1, workload_id=1, total_range=(1e-5, 1e-2), ph_range=(6.0, 9.0))
    arr = synth.host_state_arrays(t, n + 1, with_temperature=False)
    arr["mnrl_volfrac"][:, 0] = ci.vf[:, 0]
    arr["mnrl_area"][:, 0] = 1.0
    conc = ci.totals.copy()
    conc[:, 0] = ci.ph
    sp = runtime.DeviceCells.from_host(arr, n + 1)
    rc, _ = runtime.equilibrate(tab, sp, [abi.PFX_CONSTRAINT_PH] + [abi.PFX_CONSTRAINT_TOTAL] * (t.N - 1), conc,
                                synth.react_opts())
    host = sp.to_host()
    inflow = host["total"][n].copy()
    start = {k: (None if v is None else v[:n].copy()) for k, v in host.items()}

    def opts(step_dt):
        return synth.react_opts(dt=step_dt, guess_mode=abi.PFX_GUESS_PREVIOUS, max_iterations=100)

    gpu_ms = []
    for _ in range(3):
        dc = runtime.DeviceCells.from_host({k: (None if v is None else v.copy()) for k, v in start.items()}, n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = column.os_steps_device(tab, dc, col, inflow, [dt] * nsteps, opts)
        torch.cuda.synchronize()
        gpu_ms.append((time.perf_counter() - t0) * 1e3)
    out = dc.to_host()
    g = float(np.median(gpu_ms))
    res = {"workload": "100-cell 1-D column, dx 1 m, q 1 m/yr, 100 OS steps of 0.1 d, 5 primaries / 8 complexes / "
                       "kinetic calcite, Debye-Hueckel",
           "cell_steps": n * nsteps, "gpu_ms_run": g, "gpu_cell_steps_per_s": n * nsteps / (g * 1e-3),
           "newton_iterations": int(sum(st.total_iterations for st in stats)),
           "converged_every_step": bool(all(st.converged == n for st in stats)),
           "note": "launch-bound: 100 cells per launch; transport is the explicit upwind step on both sides"}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_column(t, col, start, inflow, nsteps, dt, opts, out["pri_molal"])
    log(f"column: GPU {g:.1f} ms for {nsteps} steps; {res.get('cpu_baseline')}")
    res["implicit"] = column_implicit_leg(t, tab, col, start, inflow, nsteps, dt, opts, args, log)
    return res


def column_implicit_leg(t, tab, col, start, inflow, nsteps, dt, opts, args, log):
    """The same column with the implicit transport half on the device
    (every component in one several-right-hand-side BiCGStab), the reaction
    and the volume-fraction update: eagerly (the host reads each transport
    solve's outcome), and as one captured graph per step replayed nsteps
    times back to back (no host round trip inside the run).  Both against
    the oracle's implicit column (direct transport solves)."""
    import numpy as np
    import torch

    from pflotran_amd import abi, column, runtime
    n = col.ncells
    o_t = abi.default_opts(dt=dt, porosity=col.porosity, saturation=col.saturation, volume=col.volume)
    eager_ms, graph_ms = [], []
    for _ in range(3):
        dc = runtime.DeviceCells.from_host({k: (None if v is None else v.copy()) for k, v in start.items()}, n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        column.os_steps_device_implicit(tab, dc, col, inflow, [dt] * nsteps, opts)
        torch.cuda.synchronize()
        eager_ms.append((time.perf_counter() - t0) * 1e3)
    m_eager = dc.t["pri_molal"].clone()
    f = column.grid_faces(col, inflow)
    grid = runtime.DeviceGrid(f["nx"], f["ny"], f["nz"], f["q"], f["dd"], area=f["area"], bc_q=f["bc_q"],
                              bc_dd=f["bc_dd"], bc_total=f["bc_total"])
    tr = runtime.OsTransport(grid)
    for rep in range(3):
        dg = runtime.DeviceCells.from_host({k: (None if v is None else v.copy()) for k, v in start.items()}, n)
        graph = runtime.OsStepGraph(tr, grid, tab, dg, t.N, o_t, opts(dt), rtol=1e-12, max_iterations=12)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            graph.launch()
        torch.cuda.synchronize()
        graph_ms.append((time.perf_counter() - t0) * 1e3)
        rc_g, st_g, ks_g = graph.stats()
        graph.close()
    tr.close()
    m_graph = dg.t["pri_molal"]
    out = {"gpu_ms_run_eager": float(np.median(eager_ms)), "gpu_ms_run_graph": float(np.median(graph_ms)),
           "graph_cell_steps_per_s": n * nsteps / (float(np.median(graph_ms)) * 1e-3),
           "graph_equals_eager": bool(torch.equal(m_graph, m_eager)),
           "last_step": {"reaction_converged": int(st_g.converged), "transport_reason": int(ks_g.reason),
                         "transport_iterations": int(ks_g.iterations)},
           "note": "transport: implicit, all components in one several-right-hand-side BiCGStab (rtol 1e-12); "
                   "graph: transport + reaction + volume fractions captured once, replayed per step"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        ref = orc.HostCells({k: (None if v is None else v.copy()) for k, v in start.items()}, n)
        hg = orc.HostGrid(f["nx"], f["ny"], f["nz"], f["q"], f["dd"], area=f["area"], bc_q=f["bc_q"],
                          bc_dd=f["bc_dd"], bc_total=f["bc_total"])
        acc = np.full(n, col.porosity * col.saturation * 1000.0 * col.volume / dt)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            ref.arrays["total"] = np.ascontiguousarray(orc.os_transport_step(hg, acc, ref["total"]))
            orc.react_batch(t, ref, opts(dt), nthreads=1)
            orc.update_kinetic_state(t, ref, opts(dt), nthreads=1)
        cpu_s = time.perf_counter() - t0
        m = m_graph.cpu().numpy()
        out["cpu_baseline"] = {"value": n * nsteps / cpu_s, "unit": "cell-steps/s", "cores": 1, "kind": "port",
                               "sample": "the whole run: oracle reaction step plus a direct solve of the same "
                                         "transport matrix per component",
                               "free_ion_rel_err_gpu_vs_oracle": float(np.max(np.abs(m - ref["pri_molal"])
                                                                              / ref["pri_molal"]))}
    log(f"column implicit: eager {out['gpu_ms_run_eager']:.1f} ms, graph {out['gpu_ms_run_graph']:.1f} ms for "
        f"{nsteps} steps, graph == eager: {out['graph_equals_eager']}, {out.get('cpu_baseline')}")
    return out


def cpu_baseline_column(t, col, start, inflow, nsteps, dt, opts, m_gpu):
    """The oracle's run of the column from the same start on one core
    (median of 3), and the final free ions of the device run against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    from pflotran_amd import column
    n = col.ncells
    cpu_s = []
    ref = None
    for _ in range(3):
        ref = orc.HostCells({k: (None if v is None else v.copy()) for k, v in start.items()}, n)
        t0 = time.perf_counter()
        for _ in range(nsteps):
            ref.arrays["total"] = np.ascontiguousarray(column.transport_step(col, ref["total"], inflow, dt))
            orc.react_batch(t, ref, opts(dt), nthreads=1)
            orc.update_kinetic_state(t, ref, opts(dt), nthreads=1)
        cpu_s.append(time.perf_counter() - t0)
    c = float(np.median(cpu_s))
    rel = float(np.max(np.abs(m_gpu - ref["pri_molal"]) / ref["pri_molal"]))
    return {"value": n * nsteps / c, "unit": "cell-steps/s", "cores": 1, "kind": "port", "seconds_run": c,
            "free_ion_rel_err_gpu_vs_oracle": rel}


def zones_leg(args, log):
    """Uneven batch in one persistent launch: 4 chemistry zones (N = 8, 12,
    20, 30; 3N complexes, 3 minerals) of 1M/500k/300k/200k cells, half of
    each zone within 0.05 of mineral saturation; one reaction step from the
    previous free ions.  Timed: the zoned launch with the split the
    autotuner (rebalance from measured zone times) picked, replayed as a
    HIP graph; beside it the same cells as 4 separate launches."""
    import numpy as np
    import torch

    from pflotran_amd import abi, autotune, runtime, synth, workload

    nets = workload.zone_networks()
    tabs = [runtime.Tables(t) for t in nets]
    sizes = [max(1000, int(n * args.zones_scale)) for n in workload.ZONE_CELLS]
    t0 = time.time()
    dcs, inside = [], []
    for z, (t, tab, n) in enumerate(zip(nets, tabs, sizes)):
        dc, lo, ins = workload.near_saturation_cells(t, tab, n, seed=synth.SEED + z, workload_id=3)
        dcs.append(dc)
        inside.append(float(ins.mean()))
    snaps = [workload.snapshot(dc) for dc in dcs]
    log(f"zones set-up: {sum(sizes)} cells in {time.time() - t0:.1f} s, near saturation {inside}")

    def restore():
        for dc, sn in zip(dcs, snaps):
            workload.restore(dc, sn)

    opts = synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS)
    opts.stream = torch.cuda.current_stream().cuda_stream
    # separate launches, one per zone
    restore()
    runtime.react_batch(tabs[0], dcs[0], opts)
    sep = []
    for _ in range(3):
        restore()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for tab, dc in zip(tabs, dcs):
            runtime.react_batch(tab, dc, opts)
        torch.cuda.synchronize()
        sep.append((time.perf_counter() - t1) * 1e3)
    # zoned launch, autotuned: per kernel variant (register cap) the split
    # starts from cells * N^2 and is rebalanced from the measured zone times;
    # the trials start from the step's inputs, not from the reacted state the
    # separate launches left
    restore()
    tabs, plan, lanes, lane_history = autotune.tune_zone_lanes(tabs, dcs, opts)
    history = lane_history
    best = {"variant": plan.variant, "blocks": list(plan.blocks[:4])}
    # the tuned plan eagerly and as a graph replay, alternated, so the two
    # see the same device state
    graph = runtime.ZonesGraph(tabs, dcs, opts, plan)
    g_ms, k_ms, e_ms, ek_ms = [], [], [], []
    for _ in range(5):
        restore()
        torch.cuda.synchronize()
        p2 = abi.ZonePlan()
        p2.nzones, p2.variant = plan.nzones, plan.variant
        for z in range(plan.nzones):
            p2.blocks[z] = plan.blocks[z]
        t1 = time.perf_counter()
        runtime.react_zones(tabs, dcs, opts, p2)
        torch.cuda.synchronize()
        e_ms.append((time.perf_counter() - t1) * 1e3)
        ek_ms.append(p2.kernel_ms)
        restore()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        graph.launch()
        rc, st, gp = graph.stats()
        g_ms.append((time.perf_counter() - t1) * 1e3)
        k_ms.append(gp.kernel_ms)
    graph.close()
    n = sum(sizes)
    conv = sum(s.converged for s in st)
    its = [s.total_iterations / s.ncells for s in st]
    hist = np.sum([np.array(list(s.iteration_histogram)) for s in st], axis=0)
    out = {"workload": "4 zones N=8/12/20/30 (M=3N, K=3), 1M/500k/300k/200k cells x scale, half within 0.05 "
                       "of mineral-0 saturation, previous-state guess",
           "ncells": n, "zone_cells": sizes, "near_saturation_fraction": inside,
           "graph_ms_per_step": float(np.mean(g_ms)), "graph_kernel_ms": float(np.mean(k_ms)),
           "cells_per_s": conv / (float(np.mean(g_ms)) * 1e-3),
           "separate_launches_ms": float(np.mean(sep)), "eager_tuned_plan_ms": float(np.mean(e_ms)),
           "eager_tuned_plan_kernel_ms": float(np.mean(ek_ms)), "autotune_history": history, "blocks": best["blocks"],
           "group_lanes": lanes,
           "variant": best["variant"], "graph_plan": {"variant": gp.variant, "blocks": list(gp.blocks[:4]),
                                                      "zone_ms": list(gp.zone_ms[:4])},
           "mean_iterations_per_zone": its, "iteration_histogram": hist.tolist(),
           "converged": int(conv)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_zones(nets, dcs, snaps, log)
    log(f"zones: graph {out['graph_ms_per_step']:.2f} ms/step (kernel {out['graph_kernel_ms']:.2f}), "
        f"separate {out['separate_launches_ms']:.2f} ms, its/zone {[round(x, 2) for x in its]}")
    del dcs, snaps
    torch.cuda.empty_cache()
    return out


def cpu_baseline_zones(nets, dcs, snaps, log, sample=4000):
    """Each zone's first `sample` cells by the oracle from the same start
    (the inputs restored), against the device's free ions after the graph
    replay; and the oracle's rate on every host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    from pflotran_amd import abi, synth
    threads = cpu_threads()
    out = {"cores": threads, "kind": "port", "sample_per_zone": sample, "zones": []}
    for t, dc, sn in zip(nets, dcs, snaps):
        k = min(sample, dc.ncells)
        gpu_m = dc.t["pri_molal"][:k].cpu().numpy()
        gpu_st = dc.t["status"][:k].cpu().numpy()
        arr = {key: (None if v is None else v[:k].cpu().numpy().copy()) for key, v in dc.t.items()}
        for key, v in sn.items():
            arr[key] = v[:k].cpu().numpy().copy()
        cells = orc.HostCells(arr, k)
        t0 = time.perf_counter()
        orc.react_batch(t, cells, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS), nthreads=threads)
        sec = time.perf_counter() - t0
        ok = (cells["status"] == 0) & (gpu_st == 0)
        rel = np.max(np.abs(gpu_m[ok] - cells["pri_molal"][ok]) / cells["pri_molal"][ok], axis=1)
        out["zones"].append({"N": t.N, "cells": k, "status_equal": bool(np.array_equal(gpu_st, cells["status"])),
                             "free_ion_max_rel_err": float(rel.max()), "p99": float(np.quantile(rel, 0.99)),
                             "cells_above_1e-10": int((rel > 1e-10).sum()), "oracle_cells_per_s": k / sec})
    log(f"zones vs oracle: {out}")
    return out


def cpu_threads():
    """Host cores of this process's share (affinity mask), at most 16: the
    CPU share of one GPU on the measurement machine."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(t, dc, snap, gpu_out, n, runs=5):
    """The oracle (C restatement of RReact, OpenMP over cells) on the first
    n cells of the same held-out batch (their step inputs copied to the
    host), from the previous free ions as RReact starts, timed on every core
    of this process's share and on one core (a sixteenth of the sample),
    median of `runs` runs each.  Its free ions are compared with every GPU
    method's: the same algorithm from the same start (parity) and the other
    starting points (the same root to the step test's tolerance).  Cells
    whose FP64 conditioning floor (eps times the Newton system's amplification
    of the residual's rounding) exceeds 1e-11 are counted separately: their
    free ions are not resolved to 1e-10 by FP64 itself."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as orc
    from pflotran_amd import abi, synth

    threads = cpu_threads()
    inputs = orc.HostCells.allocate(t, n, temperature=True)
    for k, v in inputs.arrays.items():
        src = snap.get(k, dc.t.get(k))
        if v is not None and src is not None:
            v[:] = src[:n].cpu().numpy().reshape(v.shape)
    opts = synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS)

    def timed(nc, nthreads):
        sub = orc.HostCells({k: (None if v is None else v[:nc].copy()) for k, v in inputs.arrays.items()}, nc)
        ts, out = [], None
        for _ in range(runs):
            c = sub.copy()
            t0 = time.perf_counter()
            rc, st = orc.react_batch(t, c, opts, nthreads=nthreads)
            ts.append(time.perf_counter() - t0)
            out = (c, st)
        return float(np.median(ts)), out

    dt_all, (cells, st) = timed(n, threads)
    n1 = max(1, n // 16)
    dt_one, (_, st1) = timed(n1, 1)
    # conditioning floor of each cell at the oracle's solution
    o = synth.react_opts()
    fixed = o.porosity * o.saturation * 1000.0 * o.volume * inputs["total"] + inputs["total_sorb"] * o.volume
    _, J = orc.residual_jacobian(t, cells.copy(), o, fixed_accum=np.ascontiguousarray(fixed), nthreads=threads)
    Jl = J * cells["pri_molal"][:, None, :]
    floor = np.full(n, np.inf)
    # each row's rounding is relative to the magnitude of its terms: the
    # fixed accumulation, phi s rho V (m + sum |nu| C) and V |S|, over dt
    absnu = np.zeros((t.M, t.N))
    for i in range(t.M):
        for s_ in range(t.cplx_specid[i, 0]):
            absnu[i, t.cplx_specid[i, 1 + s_]] += abs(t.cplx_stoich[i, s_])
    ev = cells.copy()
    orc.residual_jacobian(t, ev, o, fixed_accum=np.ascontiguousarray(fixed), nthreads=threads)  # complexes
    terms = ev["pri_molal"] + ev["sec_molal"][:, :t.M] @ absnu
    scale = (o.porosity * o.saturation * o.den_kg * o.volume * terms + np.abs(fixed)
             + np.abs(ev["total_sorb"]) * o.volume) / o.dt
    for i0 in range(0, n, 4096):
        blk = slice(i0, min(n, i0 + 4096))
        try:
            inv = np.linalg.inv(Jl[blk])
            floor[blk] = 2.220446049250313e-16 * np.max(np.abs(inv) @ scale[blk][:, :, None], axis=(1, 2))
        except np.linalg.LinAlgError:
            for i in range(blk.start, blk.stop):
                try:
                    floor[i] = 2.220446049250313e-16 * np.max(np.abs(np.linalg.inv(Jl[i])) @ scale[i])
                except np.linalg.LinAlgError:
                    pass
    resolved = floor <= 1e-11

    def compare(m_gpu, s_gpu):
        ok = (cells["status"] == 0) & (s_gpu == 0)
        if not ok.any():
            return None
        rel = np.max(np.abs(m_gpu - cells["pri_molal"]) / cells["pri_molal"], axis=1)
        rel[~ok] = 0.0
        return {"max": float(rel.max()), "max_resolved_cells": float(rel[resolved].max()),
                "p99": float(np.quantile(rel[ok], 0.99)), "median": float(np.median(rel[ok])),
                "cells": int(ok.sum()), "cells_above_1e-10": int((rel > 1e-10).sum()),
                "cells_above_1e-10_resolved": int((rel[resolved] > 1e-10).sum()),
                "max_over_8x_floor": float(np.max(rel / np.maximum(1e-10, 8.0 * floor)))}

    out = {"value": st.converged / dt_all, "unit": "cells/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": f"first {n} cells of the held-out batch, same tables and convergence test, from the previous "
                     f"free ions (RReact's start); median of {runs} runs of {dt_all:.2f} s on {threads} threads, "
                     f"{st.total_iterations} Newton iterations",
           "one_core": {"value": st1.converged / dt_one, "cells": n1, "seconds": dt_one},
           "conditioning_floor": {"cells_resolved_to_1e-11": int(resolved.sum()), "cells": n,
                                  "median": float(np.median(floor))}}
    for key, (m, s_) in gpu_out.items():
        out[f"free_ion_rel_err_{key}_gpu_vs_oracle"] = compare(m, s_)
    return out


if __name__ == "__main__":
    main()
