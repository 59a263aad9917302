"""Host side of the learned warm start (include/pfx_surrogate.h).

A multilayer perceptron (40 -> 256 -> 256 -> 256 -> 15, GELU) maps a cell's
log transported totals, its previous log free ions, temperature and mineral
volume fractions to the change of the log free-ion molalities over the
reaction step.  Its training pairs come from the GPU Newton solver itself
(speciated synthetic cells, one converged reaction step each); training and
inference run in the BF16 matrix-core kernels of libpfx_chem.so.  The
prediction only seeds the FP64 Newton (PFX_GUESS_ARRAY_FALLBACK): a cell that
does not converge from it repeats the step as the plain method does, from
the previous free ions, so every returned state meets the same convergence
test as the plain method.

PyTorch holds the device buffers and draws minibatch permutations; every
product, activation, gradient and optimizer update is a HIP kernel.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Dict, Optional

import numpy as np

from . import abi, runtime, synth, workload
from .database import ReactionTables

DIMS = (40, 256, 256, 256, 15)
TRAIN_WORKLOAD = 4            # RNG key of the training cells (held-out cells use 2)
W_WEIGHT, W_BIAS, W_GRAD, B_GRAD, W_NORM = 0, 1, 2, 3, 4


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class MLP:
    """pfx_mlp handle: FP32 master weights, Adam state and BF16 inference copy."""

    def __init__(self, dims=DIMS, batch_max: int = 8192, seed: int = synth.SEED):
        lib = runtime.load_library()
        runtime._torch()
        self.dims = tuple(int(d) for d in dims)
        self.batch_max = int(batch_max)
        d = np.asarray(self.dims, dtype=np.int32)
        h = C.c_void_p()
        rc = lib.pfx_mlp_create(len(self.dims) - 1, d.ctypes.data, self.batch_max, seed, C.byref(h))
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_create failed: {rc}")
        self.handle = h
        self.nlayers = len(self.dims) - 1

    def close(self):
        if getattr(self, "handle", None):
            runtime.load_library().pfx_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def _shape(self, which, layer):
        if which in (W_WEIGHT, W_GRAD):
            return (self.dims[layer + 1], self.dims[layer])
        if which in (W_BIAS, B_GRAD):
            return (self.dims[layer + 1],)
        return (2 * self.dims[0] + 2 * self.dims[-1],)

    def get(self, which: int, layer: int = 0) -> np.ndarray:
        import torch
        torch.cuda.synchronize()
        a = np.zeros(self._shape(which, layer), dtype=np.float32)
        rc = runtime.load_library().pfx_mlp_params(self.handle, which, layer, a.ctypes.data, 0)
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_params failed: {rc}")
        return a

    def set(self, which: int, layer: int, arr: np.ndarray):
        import torch
        torch.cuda.synchronize()
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(self._shape(which, layer))
        rc = runtime.load_library().pfx_mlp_params(self.handle, which, layer, a.ctypes.data, 1)
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_params failed: {rc}")

    def weights(self):
        return [(self.get(W_WEIGHT, l), self.get(W_BIAS, l)) for l in range(self.nlayers)]

    def gradients_host(self):
        return [(self.get(W_GRAD, l), self.get(B_GRAD, l)) for l in range(self.nlayers)]

    def get_norm(self) -> Dict[str, np.ndarray]:
        a = self.get(W_NORM)
        ni, no = self.dims[0], self.dims[-1]
        return {"in_shift": a[:ni], "in_scale": a[ni:2 * ni], "out_shift": a[2 * ni:2 * ni + no],
                "out_scale": a[2 * ni + no:]}

    def set_norm(self, in_shift, in_scale, out_shift, out_scale):
        self.set(W_NORM, 0, np.concatenate([np.asarray(v, dtype=np.float32).ravel()
                                            for v in (in_shift, in_scale, out_shift, out_scale)]))

    # ---- compute
    def forward(self, X):
        import torch
        X = X.contiguous()
        Y = torch.empty((X.shape[0], self.dims[-1]), dtype=torch.float32, device=X.device)
        rc = runtime.load_library().pfx_mlp_forward(self.handle, X.data_ptr(), X.shape[0], Y.data_ptr(), _stream())
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_forward failed: {rc}")
        return Y

    def gradients(self, X, T) -> float:
        loss = C.c_float(0.0)
        rc = runtime.load_library().pfx_mlp_gradients(self.handle, X.data_ptr(), T.data_ptr(), X.shape[0],
                                                       C.byref(loss), _stream())
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_gradients failed: {rc}")
        return float(loss.value)

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        rc = runtime.load_library().pfx_mlp_adam(self.handle, lr, beta1, beta2, eps, _stream())
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_adam failed: {rc}")

    def train_step(self, X, T, lr, want_loss=False) -> Optional[float]:
        loss = C.c_float(0.0)
        rc = runtime.load_library().pfx_mlp_train_step(self.handle, X.data_ptr(), T.data_ptr(), X.shape[0], lr,
                                                        C.byref(loss) if want_loss else None, _stream())
        if rc != 0:
            raise runtime.PfxError(f"pfx_mlp_train_step failed: {rc}")
        return float(loss.value) if want_loss else None


def features(tab: runtime.Tables, mlp: MLP, dc: runtime.DeviceCells):
    """Normalised features X[n][dims0] of the cells (device FP32)."""
    import torch
    X = torch.empty((dc.ncells, mlp.dims[0]), dtype=torch.float32, device="cuda")
    cs = dc.struct()
    rc = runtime.load_library().pfx_surrogate_features(tab.handle, mlp.handle, C.byref(cs), X.data_ptr(), _stream())
    if rc != 0:
        raise runtime.PfxError(f"pfx_surrogate_features failed: {rc}")
    return X


def guess(tab: runtime.Tables, mlp: MLP, dc: runtime.DeviceCells):
    """Fused warm start: dc['guess'] = exp(predicted ln m) (one launch)."""
    cs = dc.struct()
    rc = runtime.load_library().pfx_surrogate_guess(tab.handle, mlp.handle, C.byref(cs), _stream())
    if rc != 0:
        raise runtime.PfxError(f"pfx_surrogate_guess failed: {rc}")


def guess_from(mlp: MLP, Y, ncells: int, naqcomp: int, prev, out):
    rc = runtime.load_library().pfx_surrogate_guess_from(mlp.handle, Y.data_ptr(), ncells, naqcomp, prev.data_ptr(),
                                                          out.data_ptr(), _stream())
    if rc != 0:
        raise runtime.PfxError(f"pfx_surrogate_guess_from failed: {rc}")


def generate_pairs(t: ReactionTables, tab: runtime.Tables, mlp: MLP, n_pairs: int, chunk: int = 1_000_000,
                   seed: int = synth.SEED, workload_id: int = TRAIN_WORKLOAD, log=None):
    """(raw features, change of ln m over the step) of n_pairs synthetic
    cells, produced by the GPU solver: speciation of the previous state,
    then one reaction step from the previous free ions.  Raw = un-normalised
    features."""
    import torch
    norm = mlp.get_norm()
    ni, no = mlp.dims[0], mlp.dims[-1]
    mlp.set_norm(np.zeros(ni), np.ones(ni), np.zeros(no), np.ones(no))
    Xs, Ys = [], []
    have, k = 0, 0
    stats = {"cells": 0, "failed_speciation": 0, "not_converged": 0, "iterations": 0, "seconds": 0.0}
    t0 = time.time()
    try:
        while have < n_pairs:
            nc = min(chunk, n_pairs - have + 1024)
            dc, nfail = workload.speciated_cells(t, tab, nc, seed=seed + 7919 * k, workload_id=workload_id)
            X = features(tab, mlp, dc)
            prev = dc.t["pri_molal"].clone()
            rc, st = runtime.react_batch(tab, dc, synth.react_opts(guess_mode=abi.PFX_GUESS_PREVIOUS))
            ok = (dc.t["status"] == 0) & (dc.t["pri_molal"] > 0).all(dim=1) & (prev > 0).all(dim=1)
            idx = torch.nonzero(ok).flatten()
            Xs.append(X.index_select(0, idx))
            d = torch.log(dc.t["pri_molal"].index_select(0, idx)) - torch.log(prev.index_select(0, idx))
            Ys.append(d[:, :no].float())
            del prev
            have += int(idx.numel())
            stats["cells"] += nc
            stats["failed_speciation"] += nfail
            stats["not_converged"] += nc - int(idx.numel())
            stats["iterations"] += st.total_iterations
            k += 1
            del dc
            if log:
                log(f"pairs {have}/{n_pairs} ({time.time() - t0:.1f} s)")
    finally:
        mlp.set_norm(norm["in_shift"], norm["in_scale"], norm["out_shift"], norm["out_scale"])
    stats["seconds"] = time.time() - t0
    X = torch.cat(Xs)[:n_pairs].contiguous()
    Y = torch.cat(Ys)[:n_pairs].contiguous()
    if no > Y.shape[1]:
        Y = torch.nn.functional.pad(Y, (0, no - Y.shape[1]))
    return X, Y, stats


def fit_normalisation(mlp: MLP, X_raw, Y):
    """Per-column mean/std normalisation of features and targets, stored in
    the network; returns the normalised copies (device FP32)."""
    import torch
    mu_x = X_raw.double().mean(0)
    sd_x = X_raw.double().std(0)
    sc_x = torch.where(sd_x > 1e-6, 1.0 / sd_x, torch.ones_like(sd_x))
    mu_y = Y.double().mean(0)
    sd_y = Y.double().std(0)
    sc_y = torch.where(sd_y > 1e-6, 1.0 / sd_y, torch.ones_like(sd_y))
    in_shift, in_scale = mu_x.float(), sc_x.float()
    out_shift, out_scale = mu_y.float(), sc_y.float()
    mlp.set_norm(in_shift.cpu().numpy(), in_scale.cpu().numpy(), out_shift.cpu().numpy(), out_scale.cpu().numpy())
    Xn = ((X_raw - in_shift) * in_scale).contiguous()
    Yn = ((Y - out_shift) * out_scale).contiguous()
    return Xn, Yn


def train(mlp: MLP, Xn, Yn, epochs: float = 3.0, batch: int = 4096, lr: float = 1e-3, lr_min: float = 1e-5,
          warmup: int = 100, seed: int = synth.SEED, log=None, log_every: int = 500):
    """Adam on mean squared error of the normalised targets, cosine learning
    rate; returns a summary with the loss history."""
    import torch
    n = Xn.shape[0]
    if batch > mlp.batch_max:
        raise ValueError("batch exceeds the network's batch capacity")
    steps_per_epoch = max(1, n // batch)
    total = max(1, int(round(epochs * steps_per_epoch)))
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    hist = []
    torch.cuda.synchronize()
    t0 = time.time()
    step = 0
    perm = None
    while step < total:
        if step % steps_per_epoch == 0:
            perm = torch.randperm(n, device="cuda", generator=g)
        i = (step % steps_per_epoch) * batch
        idx = perm[i:i + batch]
        xb = Xn.index_select(0, idx)
        yb = Yn.index_select(0, idx)
        frac = step / total
        cur = lr * min(1.0, (step + 1) / warmup)
        cur = lr_min + (cur - lr_min) * 0.5 * (1.0 + math.cos(math.pi * frac))
        want = (step % log_every == 0) or step == total - 1
        loss = mlp.train_step(xb, yb, cur, want_loss=want)
        if want:
            hist.append((step, loss))
            if log:
                log(f"step {step}/{total} loss {loss:.5f} lr {cur:.2e} ({time.time() - t0:.1f} s)")
        step += 1
    torch.cuda.synchronize()
    return {"steps": total, "batch": batch, "epochs": epochs, "seconds": time.time() - t0, "loss": hist,
            "samples_per_s": total * batch / max(time.time() - t0, 1e-9)}


def prediction_error(mlp: MLP, Xn, Y_lnm):
    """RMS and max |ln m_pred - ln m| over a set of pairs, from the
    un-normalised targets (the change of ln m over the step: the same
    difference)."""
    import torch
    norm = mlp.get_norm()
    Yp = mlp.forward(Xn)
    sh = torch.as_tensor(norm["out_shift"], device="cuda")
    sc = torch.as_tensor(norm["out_scale"], device="cuda")
    lnm = Yp / sc + sh
    d = (lnm - Y_lnm)[:, :Y_lnm.shape[1]]
    return {"rms_ln": float(torch.sqrt((d * d).mean())), "max_ln": float(d.abs().max()),
            "p99_ln": float(torch.quantile(d.abs().flatten()[:1_000_000], 0.99))}
