"""Autotuning of the per-cell kernels by measurement.

The knobs are the kernel variant of pfx_tables_tune (the register cap: how
many wavefronts per SIMD the allocation must allow, trading spills for
latency hiding) and, for zoned launches, the variant, the split of
workgroups between zones (pfx_zone_plan) and each zone's group width (the
lanes per cell, the LU's row tile: pfx_reaction_desc.group_lanes).  Each
candidate runs the real step on the real cells, with the step's inputs
restored between trials, and the fastest device time wins.
"""
from __future__ import annotations

import dataclasses

from . import abi, runtime, workload


def tune_kernel(tab: runtime.Tables, dc: runtime.DeviceCells, run, kernel: int = abi.PFX_KERNEL_REACT,
                variants=(0, 1, 2), reps: int = 2):
    """Time `run()` (one launch of `kernel` on dc) for each variant; keep the
    fastest.  Returns {variant: best kernel ms}."""
    snap = workload.snapshot(dc)
    times = {}
    for v in variants:
        tab.tune(kernel, v)
        best = None
        for _ in range(reps):
            workload.restore(dc, snap)
            run()
            ms = tab.timing()[0]
            best = ms if best is None else min(best, ms)
        times[v] = best
    workload.restore(dc, snap)
    tab.tune(kernel, min(times, key=times.get))
    return times


def tune_zones(tabs, dcs, opts: abi.ReactOpts, variants=(0, 1, 2), rounds: int = 3):
    """Zoned launch: for each kernel variant, start from the cells * N^2
    split and rebalance from the measured zone times; return the fastest
    plan and the history of every trial."""
    snaps = [workload.snapshot(dc) for dc in dcs]
    history = []
    best = None
    for variant in variants:
        plan = abi.ZonePlan()
        plan.nzones = len(tabs)
        plan.variant = variant
        for _ in range(rounds):
            for dc, sn in zip(dcs, snaps):
                workload.restore(dc, sn)
            rc, st, plan = runtime.react_zones(tabs, dcs, opts, plan)
            rec = {"variant": variant, "blocks": list(plan.blocks[:plan.nzones]),
                   "zone_ms": list(plan.zone_ms[:plan.nzones]), "kernel_ms": plan.kernel_ms}
            history.append(rec)
            if best is None or rec["kernel_ms"] < best["kernel_ms"]:
                best = rec
            runtime.rebalance(plan)
    for dc, sn in zip(dcs, snaps):
        workload.restore(dc, sn)
    out = abi.ZonePlan()
    out.nzones = len(tabs)
    out.variant = best["variant"]
    for z, b in enumerate(best["blocks"]):
        out.blocks[z] = b
    return out, history


def default_lanes(N: int) -> int:
    return 8 if N <= 8 else (16 if N <= 16 else 32)


def tune_zone_lanes(tabs, dcs, opts: abi.ReactOpts, rounds: int = 3):
    """The zoned launch with each zone's group width chosen by measurement:
    the default width (the smallest of 8/16/32 holding N) against twice it
    where that fits (<= 32), one zone at a time, each trial with the split
    and register variant tuned by tune_zones.  Returns (the tables to use,
    one per zone; the plan; the lanes per zone; the trial history)."""
    plan, hist = tune_zones(tabs, dcs, opts, rounds=rounds)
    best_ms = min(h["kernel_ms"] for h in hist)
    lanes = [default_lanes(t.host.N) for t in tabs]
    chosen = list(tabs)
    history = [{"lanes": list(lanes), "kernel_ms": best_ms}]
    for z, t in enumerate(tabs):
        wide = 2 * lanes[z]
        if wide > 32:
            continue
        alt = runtime.Tables(dataclasses.replace(t.host, group_lanes=wide))
        trial = list(chosen)
        trial[z] = alt
        p2, h2 = tune_zones(trial, dcs, opts, variants=(plan.variant,), rounds=rounds)
        ms = min(h["kernel_ms"] for h in h2)
        trial_lanes = list(lanes)
        trial_lanes[z] = wide
        history.append({"lanes": trial_lanes, "kernel_ms": ms})
        if ms < best_ms:
            best_ms, plan, chosen, lanes = ms, p2, trial, trial_lanes
    return chosen, plan, lanes, history
