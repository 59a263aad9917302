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
  from the default guess, same convergence, same free ions (1e-10).
"""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as orc
from gpu_cases import host_arrays, speciated_state
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
    Xo = mo.features(cells["total"], cells["total_sorb"], cells["temperature"], cells["mnrl_volfrac"], 40, t.N,
                     t.K, t.R)
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
    Xo = mo.features(cells["total"], cells["total_sorb"], cells["temperature"], cells["mnrl_volfrac"], 40, t.N,
                     t.K, t.R)
    sd = Xo.std(0)
    in_scale = np.where(sd > 1e-6, 1.0 / np.where(sd > 1e-6, sd, 1.0), 1.0)
    out_shift = rng.uniform(-20, -5, size=15)
    out_scale = rng.uniform(0.1, 1.0, size=15)
    mlp.set_norm(Xo.mean(0), in_scale, out_shift, out_scale)
    surrogate.guess(tab, mlp, dc)
    g_fused = dc.t["guess"].cpu().numpy()
    X = surrogate.features(tab, mlp, dc)
    Y = mlp.forward(X)
    g2 = torch.zeros((n, t.N), dtype=torch.float64, device="cuda")
    surrogate.guess_from(mlp, Y, n, t.N, g2)
    g_unfused = g2.cpu().numpy()
    nrm = mlp.get_norm()
    y_fused = (np.log(g_fused) - nrm["out_shift"]) * nrm["out_scale"]
    y_unfused = (np.log(g_unfused) - nrm["out_shift"]) * nrm["out_scale"]
    Yq, _ = mo.forward(mlp.weights(), X.cpu().numpy(), bf16=True)
    assert np.all(np.isfinite(g_fused)) and np.all(g_fused > 0)
    assert rel_to_max(y_fused, y_unfused) <= 2e-3
    assert rel_to_max(y_fused, Yq) <= 2e-3


def test_fallback_is_the_plain_method_when_every_warm_attempt_fails(torch, net15):
    t = net15
    n = 300
    cells, _, _, _ = speciated_state(t, n)
    tab = runtime.Tables(t)
    arr = host_arrays(cells)
    arr["guess"][:] = 1e3   # far from every solution: no cell converges within 3 iterations
    dc = runtime.DeviceCells.from_host(arr, n)
    snap = workload.snapshot(dc)
    rc, st_fb = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK,
                                                              fallback_after=3, extra_iterations=1))
    fb = {k: dc.t[k].cpu().numpy() for k in ("pri_molal", "num_iterations", "status", "total")}
    workload.restore(dc, snap)
    rc, st_pl = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT, extra_iterations=1))
    pl = {k: dc.t[k].cpu().numpy() for k in ("pri_molal", "num_iterations", "status", "total")}
    assert st_fb.fallbacks == n
    assert np.array_equal(fb["pri_molal"], pl["pri_molal"])
    assert np.array_equal(fb["total"], pl["total"])
    assert np.array_equal(fb["status"], pl["status"])
    extra = fb["num_iterations"] - pl["num_iterations"]   # iterations of the abandoned attempt
    assert np.all((extra >= 1) & (extra <= 3))
    # the oracle's fallback takes the same path
    oc = cells.copy()
    oc["guess"][:] = 1e3
    rc, st_o = orc.react_batch(t, oc, synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK, fallback_after=3,
                                                       extra_iterations=1))
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
    t = net15
    tab = runtime.Tables(t)
    mlp = surrogate.MLP(surrogate.DIMS, batch_max=2048, seed=123)
    X_raw, Y, gst = surrogate.generate_pairs(t, tab, mlp, 120_000, chunk=60_000)
    Xn, Yn = surrogate.fit_normalisation(mlp, X_raw, Y)
    tr = surrogate.train(mlp, Xn, Yn, epochs=4, batch=2048, lr=2e-3)
    assert tr["loss"][-1][1] < 0.5 * tr["loss"][0][1]
    # held-out cells of the bench workload (RNG key 2, not the training key)
    n = 4000
    dc, _ = workload.speciated_cells(t, tab, n)
    snap = workload.snapshot(dc)
    rc, st_plain = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_DEFAULT,
                                                                 extra_iterations=1))
    m_plain = dc.t["pri_molal"].cpu().numpy()
    s_plain = dc.t["status"].cpu().numpy()
    workload.restore(dc, snap)
    surrogate.guess(tab, mlp, dc)
    rc, st_warm = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_ARRAY_FALLBACK,
                                                                fallback_after=30, extra_iterations=1))
    m_warm = dc.t["pri_molal"].cpu().numpy()
    s_warm = dc.t["status"].cpu().numpy()
    print(f"plain {st_plain.total_iterations / n:.2f} it/cell, warm {st_warm.total_iterations / n:.2f} it/cell, "
          f"fallbacks {st_warm.fallbacks}, converged {st_plain.converged}/{st_warm.converged}")
    assert st_warm.converged >= st_plain.converged
    assert st_warm.total_iterations < 0.8 * st_plain.total_iterations
    # every returned state passes the same FP64 acceptance test
    rr_warm = dc.t["rel_res"].cpu().numpy()
    assert np.all(rr_warm[s_warm == 0] <= 1e-10)
    # both methods reach the same free ions; cells whose Newton converges
    # only linearly (capped mineral lnQK in the Jacobian) stop at different
    # iterates within the tolerance
    both = (s_plain == 0) & (s_warm == 0)
    assert both.sum() >= 0.99 * n
    rel = np.max(np.abs(m_warm[both] - m_plain[both]) / m_plain[both], axis=1)
    print(f"free-ion difference: median {np.median(rel):.2e}, 99% {np.quantile(rel, 0.99):.2e}, max {rel.max():.2e}")
    assert np.quantile(rel, 0.99) <= 1e-10
    assert rel.max() <= 1e-6
