"""Learned warm start on the GPU, through the C-ABI:

* the BF16 matrix-core forward pass, gradients and Adam step against the
  FP64 oracle (oracle/mlp_oracle.py) with the same BF16 operand rounding
  (tolerance 2e-3 of the largest entry; an operand-layout mistake gives
  O(1) errors) and against the unrounded FP64 network (3e-2, the BF16
  budget);
* the feature kernel against the oracle's feature map (1e-6 relative);
* the fused inference kernel against the unfused path and the oracle;
* the fallback guess mode: bit-identical to the plain method when every
  warm attempt fails, and equal to the CPU oracle's fallback;
* a trained network's gain: fewer Newton iterations than the plain method
  from the previous free ions (RReact's own start) and from the default
  guess, same convergence, and free ions of every method against the
  oracle to 1e-10 over 20,000 cells of the timed workload (the largest
  error, not a percentile), wherever FP64 resolves the free ions that
  finely (8x the per-cell conditioning floor elsewhere).
"""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as orc
from gpu_cases import conditioning_floor, host_arrays, speciated_state
from pflotran_amd import abi, runtime, surrogate, synth, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    return _t


@pytest.fixture(scope="module")
def net15():
    return synth.random_network(15, 50, 3, ionx=True)


def rel_to_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def make_mlp(batch_max=4200, seed=5):
    return surrogate.MLP(surrogate.DIMS, batch_max=batch_max, seed=seed)


def perturb_biases(mlp, rng):
    # non-zero biases so the bias path is exercised
    for l in range(mlp.nlayers):
        mlp.set(surrogate.W_BIAS, l, rng.normal(0, 0.3, size=mlp.dims[l + 1]))


@pytest.mark.parametrize("B", [1, 300])
def test_forward_matches_oracle(torch, B):
    rng = np.random.default_rng(B)
    mlp = make_mlp()
    perturb_biases(mlp, rng)
    X = rng.normal(size=(B, 40)).astype(np.float32)
    Y = mlp.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    params = mlp.weights()
    Yq, _ = mo.forward(params, X, bf16=True)
    Y64, _ = mo.forward(params, X, bf16=False)
    assert rel_to_max(Y, Yq) <= 2e-3
    assert rel_to_max(Y, Y64) <= 3e-2


@pytest.mark.parametrize("B", [300, 4100])
def test_gradients_match_oracle(torch, B):
    # B = 4100 takes the split-K (atomic) path of the weight gradients
    rng = np.random.default_rng(B + 1)
    mlp = make_mlp()
    perturb_biases(mlp, rng)
    X = rng.normal(size=(B, 40)).astype(np.float32)
    T = rng.normal(size=(B, 15)).astype(np.float32)
    loss = mlp.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    grads = mlp.gradients_host()
    params = mlp.weights()
    loss_q, grads_q = mo.loss_and_grads(params, X, T, bf16=True)
    loss_64, grads_64 = mo.loss_and_grads(params, X, T, bf16=False)
    assert abs(loss - loss_q) <= 1e-3 * loss_q
    for l in range(mlp.nlayers):
        for g, gq, g64 in zip(grads[l], grads_q[l], grads_64[l]):
            assert rel_to_max(g, gq) <= 2e-3, l
            assert rel_to_max(g, g64) <= 3e-2, l


def test_adam_matches_oracle(torch):
    rng = np.random.default_rng(7)
    mlp = make_mlp(batch_max=64)
    w0 = mlp.weights()
    gs = []
    for l in range(mlp.nlayers):
        gw = rng.normal(size=w0[l][0].shape).astype(np.float32)
        gb = rng.normal(size=w0[l][1].shape).astype(np.float32)
        mlp.set(surrogate.W_GRAD, l, gw)
        mlp.set(surrogate.B_GRAD, l, gb)
        gs.append((gw, gb))
    for step in (1, 2):
        mlp.adam(1e-3)
    torch.cuda.synchronize()
    w2 = mlp.weights()
    for l in range(mlp.nlayers):
        for k in range(2):
            w, m, v = w0[l][k].astype(np.float64), 0.0, 0.0
            for step in (1, 2):
                w, m, v = mo.adam_step(w, gs[l][k].astype(np.float64), m, v, step, 1e-3)
            assert np.max(np.abs(w2[l][k] - w)) <= 1e-6, (l, k)


def test_features_match_oracle(torch, net15):
    t = net15
    n = 97
    cells, _, _, _ = speciated_state(t, n)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    mlp = make_mlp(batch_max=64)
    X = surrogate.features(tab, mlp, dc).cpu().numpy()
    Xo = mo.features(cells["total"], cells["pri_molal"], cells["temperature"], cells["mnrl_volfrac"], 40, t.N,
                     t.K)
    assert np.max(np.abs(X - Xo) / np.maximum(np.abs(Xo), 1.0)) <= 1e-6


def test_fused_guess_matches_unfused_and_oracle(torch, net15):
    t = net15
    n = 1000   # 15 full blocks of 64 cells and a ragged one
    cells, _, _, _ = speciated_state(t, n)
    tab = runtime.Tables(t)
    dc = runtime.DeviceCells.from_host(host_arrays(cells), n)
    rng = np.random.default_rng(9)
    mlp = make_mlp(batch_max=512)
    perturb_biases(mlp, rng)
    Xo = mo.features(cells["total"], cells["pri_molal"], cells["temperature"], cells["mnrl_volfrac"], 40, t.N,
                     t.K)
    sd = Xo.std(0)
    in_scale = np.where(sd > 1e-6, 1.0 / np.where(sd > 1e-6, sd, 1.0), 1.0)
    # outputs are changes of ln m: kept well inside the guess's +-30 hold
    out_shift = rng.uniform(-2.0, 2.0, size=15)
    out_scale = rng.uniform(0.5, 2.0, size=15)
    mlp.set_norm(Xo.mean(0), in_scale, out_shift, out_scale)
    surrogate.guess(tab, mlp, dc)
    g_fused = dc.t["guess"].cpu().numpy()
    X = surrogate.features(tab, mlp, dc)
    Y = mlp.forward(X)
    g2 = torch.zeros((n, t.N), dtype=torch.float64, device="cuda")
    surrogate.guess_from(mlp, Y, n, t.N, dc.t["pri_molal"], g2)
    g_unfused = g2.cpu().numpy()
    nrm = mlp.get_norm()
    prev = np.log(cells["pri_molal"])
    y_fused = (np.log(g_fused) - prev - nrm["out_shift"]) * nrm["out_scale"]
    y_unfused = (np.log(g_unfused) - prev - nrm["out_shift"]) * nrm["out_scale"]
    Yq, _ = mo.forward(mlp.weights(), X.cpu().numpy(), bf16=True)
    assert np.all(np.isfinite(g_fused)) and np.all(g_fused > 0)
    assert rel_to_max(y_fused, y_unfused) <= 2e-3
    assert rel_to_max(y_fused, Yq) <= 2e-3


@pytest.mark.parametrize("act_freq,fallback_previous", [(abi.PFX_ACT_TIMESTEP, 0), (abi.PFX_ACT_NEWTON_ITER, 0),
                                                          (abi.PFX_ACT_TIMESTEP, 1)])
def test_fallback_is_the_plain_method_when_every_warm_attempt_fails(torch, act_freq, fallback_previous):
    t = synth.random_network(15, 50, 3, ionx=True, act_freq=act_freq)
    n = 300
    cells, _, _, _ = speciated_state(t, n)
    tab = runtime.Tables(t)
    arr = host_arrays(cells)
    arr["guess"][:] = 1e3   # far from every solution: no cell converges within 3 iterations
    dc = runtime.DeviceCells.from_host(arr, n)
    snap = workload.snapshot(dc)
    plain_mode = abi.PFX_GUESS_PREVIOUS if fallback_previous else abi.PFX_GUESS_DEFAULT
    fb_opts = dict(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK, fallback_after=3, fallback_previous=fallback_previous)
    rc, st_fb = runtime.react_batch(tab, dc, synth.react_opts(**fb_opts))
    fb = {k: dc.t[k].cpu().numpy() for k in ("pri_molal", "num_iterations", "status", "total", "pri_act_coef",
                                             "sec_act_coef", "sec_molal")}
    workload.restore(dc, snap)
    rc, st_pl = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=plain_mode))
    pl = {k: dc.t[k].cpu().numpy() for k in fb}
    assert st_fb.fallbacks == n
    # the restart repeats the plain method from the step's start: activity
    # coefficients, complexes and free ions bit for bit (with activity every
    # iteration from the 1e-9 default guess, cells that end non-finite end
    # so in both)
    for k in ("pri_molal", "total", "status", "pri_act_coef", "sec_act_coef", "sec_molal"):
        assert np.array_equal(fb[k], pl[k], equal_nan=fb[k].dtype.kind == "f"), k
    # iterations of the abandoned attempt: up to fallback_after; none where
    # the attempt turns non-finite at its first evaluation (with activity
    # every iteration, the ionic strength at 1e3 molal does so in every cell)
    extra = fb["num_iterations"] - pl["num_iterations"]
    assert np.all((extra >= 0) & (extra <= 3))
    if act_freq == abi.PFX_ACT_TIMESTEP:
        assert np.mean(extra >= 1) > 0.5
    # the oracle's fallback takes the same path
    oc = cells.copy()
    oc["guess"][:] = 1e3
    rc, st_o = orc.react_batch(t, oc, synth.react_opts(**fb_opts))
    assert st_o.fallbacks == n
    # iteration counts agree closely (rounding can move a long clipped path)
    assert np.mean(np.abs(oc["num_iterations"] - fb["num_iterations"])) < 0.5
    ok = oc["status"] == 0
    assert np.array_equal(oc["status"], fb["status"])
    assert np.max(np.abs(fb["pri_molal"][ok] - oc["pri_molal"][ok]) / oc["pri_molal"][ok]) <= 1e-10


def test_fallback_untouched_by_a_good_guess(torch, net15):
    t = net15
    n = 200
    cells, _, _, _ = speciated_state(t, n)
    tab = runtime.Tables(t)
    # the converged state of the step, perturbed by 1%, as the guess
    ref = cells.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT, extra_iterations=1))
    arr = host_arrays(cells)
    arr["guess"][:] = ref["pri_molal"] * np.exp(0.01 * np.sin(np.arange(n * t.N)).reshape(n, t.N))
    dc = runtime.DeviceCells.from_host(arr, n)
    # the warm attempt gets the whole budget: a few cells converge slowly even
    # from 1% off the solution (the reference's Jacobian omits the capped
    # mineral lnQK and uses den/1000 in the mineral terms), but none fails
    rc, st = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK,
                                                           extra_iterations=1))
    conv = ref["status"] == 0
    its = dc.t["num_iterations"].cpu().numpy()
    assert st.fallbacks <= int((~conv).sum())
    assert st.converged >= int(conv.sum())
    assert np.median(its[conv]) <= 5
    assert np.mean(its[conv]) < np.mean(ref["num_iterations"][conv]) / 2
    m = dc.t["pri_molal"].cpu().numpy()
    assert np.max(np.abs(m[conv] - ref["pri_molal"][conv]) / ref["pri_molal"][conv]) <= 1e-10


def test_trained_network_reduces_newton_iterations(torch, net15):
    """A network trained on GPU-generated pairs against both plain starting
    points, on 20,000 held-out cells of the 1M-cell workload, with the timed
    path's options: fewer Newton iterations than RReact's own start (the
    previous free ions) and than the default guess; every cell converged by
    every method; free ions of all three against the oracle's (default
    guess) to 1e-10 wherever FP64 resolves them that finely."""
    t = net15
    tab = runtime.Tables(t)
    mlp = surrogate.MLP(surrogate.DIMS, batch_max=2048, seed=123)
    X_raw, Y, gst = surrogate.generate_pairs(t, tab, mlp, 200_000, chunk=100_000)
    Xn, Yn = surrogate.fit_normalisation(mlp, X_raw, Y)
    tr = surrogate.train(mlp, Xn, Yn, epochs=6, batch=2048, lr=2e-3)
    assert tr["loss"][-1][1] < 0.5 * tr["loss"][0][1]
    # held-out cells of the bench workload (RNG key 2, not the training key)
    n = 20_000
    dc, _ = workload.speciated_cells(t, tab, n)
    snap = workload.snapshot(dc)
    res = {}
    for key, opts in (("default", synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT)),
                      ("previous", synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS)),
                      ("warm", synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK, fallback_after=50,
                                                fallback_previous=1))):
        workload.restore(dc, snap)
        if key == "warm":
            surrogate.guess(tab, mlp, dc)
        rc, st = runtime.react_batch(tab, dc, opts)
        res[key] = (st, dc.t["pri_molal"].cpu().numpy(), dc.t["status"].cpu().numpy(), dc.t["rel_res"].cpu().numpy())
        print(f"{key}: {st.total_iterations / n:.3f} it/cell, converged {st.converged}, fallbacks {st.fallbacks}")
    st_d, st_p, st_w = res["default"][0], res["previous"][0], res["warm"][0]
    assert st_w.converged == st_p.converged == n
    assert st_w.total_iterations < 0.9 * st_p.total_iterations
    assert st_w.total_iterations < 0.5 * st_d.total_iterations
    # every returned state passes the same FP64 acceptance test
    assert np.all(res["warm"][3] <= 1e-10)
    # the oracle from the default guess on the same cells
    workload.restore(dc, snap)
    inputs = orc.HostCells({k: (None if v is None else v.cpu().numpy().copy()) for k, v in dc.t.items()}, n)
    ref = inputs.copy()
    orc.react_batch(t, ref, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT))
    floor = conditioning_floor(t, ref, inputs, synth.STEP_DT)
    print(f"conditioning floor: median {np.median(floor):.1e}, cells above 1e-11: {(floor > 1e-11).sum()}")
    for key in ("default", "previous", "warm"):
        m, s_ = res[key][1], res[key][2]
        both = (ref["status"] == 0) & (s_ == 0)
        assert both.sum() >= n - 10
        rel = np.max(np.abs(m - ref["pri_molal"]) / ref["pri_molal"], axis=1)
        rel[~both] = 0.0
        bound = np.maximum(1e-10, 8.0 * floor)
        print(f"{key} vs oracle: max {rel.max():.2e}, over cells resolved to 1e-11: {rel[floor <= 1e-11].max():.2e}, "
              f"cells above 1e-10: {(rel > 1e-10).sum()}")
        worst = int(np.argmax(rel / bound))
        assert rel[worst] <= bound[worst], (key, worst, rel[worst], floor[worst])
        assert rel[floor <= 1e-11].max() <= 1e-10
