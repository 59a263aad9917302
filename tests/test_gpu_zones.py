"""Chemistry zones in one persistent launch (pfx_react_zones) and its HIP
graph: the zoned launch gives every zone the same bits as its own
pfx_react_batch; graph replays, with eager work and synchronisation between
them, give the same bits as the eager zoned launch; the 32-lane group (N up
to 32) matches the oracle; the near-saturation workload lands where it is
asked to."""
import numpy as np
import pytest

import oracle as orc
from gpu_cases import host_arrays, speciated_state
from pflotran_amd import abi, runtime, synth, workload

pytestmark = pytest.mark.gpu

FIELDS = ("pri_molal", "total", "pri_act_coef", "sec_molal", "mnrl_rate", "num_iterations", "status", "rel_res")


@pytest.fixture(scope="module")
def zones():
    nets = workload.zone_networks()
    sizes = (300, 200, 150, 100)
    cells = [speciated_state(t, n, seed=31 + i)[0] for i, (t, n) in enumerate(zip(nets, sizes))]
    return nets, cells


def device(cells, n):
    return runtime.DeviceCells.from_host(host_arrays(cells), n)


def outputs(dc):
    return {k: dc.t[k].cpu().numpy().copy() for k in FIELDS if dc.t.get(k) is not None}


def test_zoned_launch_equals_per_zone_batches(zones):
    nets, cells = zones
    tabs = [runtime.Tables(t) for t in nets]
    opts = lambda: synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT)  # noqa: E731
    ref = []
    for tab, c in zip(tabs, cells):
        dc = device(c, c.ncells)
        runtime.react_batch(tab, dc, opts())
        ref.append(outputs(dc))
    dcs = [device(c, c.ncells) for c in cells]
    rc, stats, plan = runtime.react_zones(tabs, dcs, opts())
    assert plan.nzones == 4 and all(plan.blocks[z] >= 1 for z in range(4))
    for z, dc in enumerate(dcs):
        out = outputs(dc)
        for k in FIELDS:
            if k in out:
                assert np.array_equal(out[k], ref[z][k]), (z, k)
        assert stats[z].ncells == cells[z].ncells
        assert stats[z].converged == int(np.sum(ref[z]["status"] == 0))
    # the rebalanced split keeps the total and gives every zone a workgroup
    total = sum(plan.blocks[z] for z in range(4))
    runtime.rebalance(plan)
    assert sum(plan.blocks[z] for z in range(4)) == total
    assert all(plan.blocks[z] >= 1 for z in range(4))


def test_zones_graph_replays_match_eager(zones):
    import torch
    nets, cells = zones
    tabs = [runtime.Tables(t) for t in nets]
    dcs = [device(c, c.ncells) for c in cells]
    snaps = [workload.snapshot(dc) for dc in dcs]
    opts = synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT)
    opts.stream = torch.cuda.current_stream().cuda_stream
    rc, st_e, _ = runtime.react_zones(tabs, dcs, opts)
    eager = [outputs(dc) for dc in dcs]
    graph = runtime.ZonesGraph(tabs, dcs, opts)
    for replay in range(3):
        for dc, sn in zip(dcs, snaps):
            workload.restore(dc, sn)
            dc.t["pri_molal"].mul_(1.0)       # eager work between replays
        torch.cuda.synchronize()
        graph.launch()
        rc_g, st_g, plan = graph.stats()
        torch.cuda.synchronize()
        for z, dc in enumerate(dcs):
            out = outputs(dc)
            for k in FIELDS:
                if k in out:
                    assert np.array_equal(out[k], eager[z][k]), (replay, z, k)
            assert st_g[z].converged == st_e[z].converged
            assert st_g[z].total_iterations == st_e[z].total_iterations
    graph.close()


def test_thirty_primaries_match_oracle(zones):
    # N = 30 runs in 32-lane groups
    nets, cells = zones
    t, c = nets[3], cells[3]
    ref = c.copy()
    orc.react_batch(t, ref, synth.react_opts())
    tab = runtime.Tables(t)
    dc = device(c, c.ncells)
    runtime.react_batch(tab, dc, synth.react_opts())
    out = dc.to_host()
    ok = ref["status"] == 0
    assert np.array_equal(out["status"], ref["status"])
    assert np.max(np.abs(out["pri_molal"][ok] - ref["pri_molal"][ok]) / ref["pri_molal"][ok]) <= 1e-10


def test_near_saturation_workload():
    t = workload.zone_networks()[0]
    tab = runtime.Tables(t)
    dc, lo, inside = workload.near_saturation_cells(t, tab, 2000, seed=synth.SEED, workload_id=3)
    assert dc.ncells == 2000
    # exactly half of the cells sit inside the window after the refresh
    assert inside.sum() == 1000
    assert np.all(np.abs(lo[inside]) <= 0.05)


def test_kernel_variants_give_identical_bits(zones):
    # the register cap changes register allocation and spills, not arithmetic
    nets, cells = zones
    for t, c in ((nets[1], cells[1]), (nets[3], cells[3])):
        tab = runtime.Tables(t)
        outs = []
        for v in (0, 1, 2):
            assert tab.tune(abi.PFX_KERNEL_REACT, v) == v
            assert tab.tune(abi.PFX_KERNEL_REACT) == v
            dc = device(c, c.ncells)
            runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT))
            outs.append(outputs(dc))
        for o in outs[1:]:
            for k in FIELDS:
                if k in o:
                    assert np.array_equal(o[k], outs[0][k]), k
    with pytest.raises(runtime.PfxError):
        tab.tune(abi.PFX_KERNEL_EVAL, 1)


def test_wider_cell_groups_give_the_same_step(zones):
    # the group width (lanes per cell, the LU's row tile) is a tuning knob:
    # padding lanes carry identity rows and add zeros to the group sums, so
    # a network run 16 or 32 lanes wide takes the same Newton path
    import dataclasses
    nets, cells = zones
    for z in (0, 1):
        t = nets[z]
        base = runtime.Tables(t)
        dc = device(cells[z], cells[z].ncells)
        runtime.react_batch(base, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
        ref = outputs(dc)
        wide = 2 * (8 if t.N <= 8 else 16)
        tw = runtime.Tables(dataclasses.replace(t, group_lanes=wide))
        dw = device(cells[z], cells[z].ncells)
        runtime.react_batch(tw, dw, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
        out = outputs(dw)
        assert np.array_equal(out["status"], ref["status"])
        ok = ref["status"] == 0
        rel = np.abs(out["pri_molal"][ok] - ref["pri_molal"][ok]) / ref["pri_molal"][ok]
        assert rel.max() <= 1e-12, (z, wide, rel.max())
    # a width below N, or not 8/16/32, is refused
    with pytest.raises(runtime.PfxError):
        runtime.Tables(dataclasses.replace(nets[1], group_lanes=8))
