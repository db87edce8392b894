"""The early-stop network: definition, training and weight packing.

Architecture (fixed; the HIP inference kernels in csrc/earlystop_net.hpp implement
exactly this):
  input [2, 375]  (V and I prefix, mean-pooled by 4, each scaled by its max |x|)
  4 x Conv1d(k=5, stride=2, pad=2) + ReLU, 64 channels: lengths 375 -> 188 -> 94 -> 47 -> 24
  flatten 1536 -> Linear 512 + ReLU -> Linear 1375
  output: 2 x 687 pooled tail samples (V then I, each in units of the prefix's final
          amplitude: target = tail / end_amplitude) and an estimate of log10 of the
          residual energy -- the energy of the difference between the predicted and the
          true tail -- over the run's total port energy (prefix plus tail, in the
          prefix-normalised units of the input).  The fallback to the full-length run
          keys on that estimate.
Weights are BF16, activations and accumulation FP32 (training uses torch autocast in
bfloat16 for the matrix products).

    python -m fdtdhip.earlystop_net train DATA.npz [DATA2.npz ...] --out weights.npz
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import earlystop as ES

CH, KW, HIDDEN = 64, 5, 512
LENS = [ES.N_IN, 188, 94, 47, 24]
N_OUT = 2 * ES.N_OUT + 1
END_WINDOW = 16          # pooled input samples that define the prefix's final amplitude


def end_amplitude(x: np.ndarray) -> np.ndarray:
    """[..., 2, N_IN] -> [..., 2]: max |x| over the last END_WINDOW pooled samples."""
    a = np.max(np.abs(x[..., -END_WINDOW:]), axis=-1)
    return np.maximum(a, 1e-12)


def make_model():
    import torch.nn as nn
    layers = []
    cin = 2
    for _ in range(4):
        layers += [nn.Conv1d(cin, CH, KW, stride=2, padding=KW // 2), nn.ReLU()]
        cin = CH
    layers += [nn.Flatten(), nn.Linear(CH * LENS[-1], HIDDEN), nn.ReLU(), nn.Linear(HIDDEN, N_OUT)]
    return nn.Sequential(*layers)


def energies(x: np.ndarray, tail: np.ndarray):
    """(prefix energy, tail energy) in normalised units from pooled samples: each
    input sample stands for POOL_IN steps, each tail sample for POOL_OUT."""
    return ES.POOL_IN * np.sum(x ** 2, axis=(-2, -1)), ES.POOL_OUT * np.sum(tail ** 2, axis=(-2, -1))


def load_data(paths):
    """x [n,2,375], y [n,1374] (scaled tail), a [n,2] end amplitudes, e_total [n]."""
    xs, ys, amps, etot = [], [], [], []
    for p in paths:
        d = np.load(p)
        x, tail = d["x"], d["tail"]
        a = end_amplitude(x)
        ep, et = energies(x, tail)
        xs.append(x)
        ys.append((tail / a[:, :, None]).reshape(len(x), -1))
        amps.append(a)
        etot.append(ep + et)
    return (np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32),
            np.concatenate(amps).astype(np.float32), np.concatenate(etot).astype(np.float32))


def residual_log_ratio(pred_tail, true_tail, a, e_total):
    """log10 of the pooled-tail error energy over the total energy (torch tensors)."""
    import torch
    err = ((pred_tail - true_tail).view(-1, 2, ES.N_OUT) * a[:, :, None]) ** 2
    return torch.log10(ES.POOL_OUT * err.sum(dim=(1, 2)) / e_total + 1e-30)


def train(paths, out, epochs=300, batch=64, lr=2e-3, val_frac=0.1, seed=0, threads=8):
    import torch
    torch.manual_seed(seed)
    torch.set_num_threads(threads)
    x, y, a, et = load_data(paths)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(x))
    nval = max(1, int(len(x) * val_frac))
    tr, va = perm[nval:], perm[:nval]
    X, Y, A, ET = torch.from_numpy(x), torch.from_numpy(y), torch.from_numpy(a), torch.from_numpy(et)
    n_tail = 2 * ES.N_OUT
    model = make_model()

    def losses(idx, pred):
        pt = pred[:, :n_tail].float()
        tail_loss = ((pt - Y[idx]) ** 2).mean()
        # The last output learns the (detached) residual energy of the current prediction.
        target = residual_log_ratio(pt.detach(), Y[idx], A[idx], ET[idx])
        est_loss = ((pred[:, -1].float() - target) ** 2).mean()
        return tail_loss, est_loss
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=1e-5)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=epochs * ((len(tr) + batch - 1) // batch))
    t0 = time.time()
    hist = []
    for ep in range(epochs):
        model.train()
        order = tr[rng.permutation(len(tr))]
        for k in range(0, len(order), batch):
            idx = torch.from_numpy(order[k:k + batch])
            with torch.autocast("cpu", dtype=torch.bfloat16):
                pred = model(X[idx])
            tl, el = losses(idx, pred)
            loss = tl + 1e-3 * el
            opt.zero_grad()
            loss.backward()
            opt.step()
            sched.step()
        if ep % 20 == 0 or ep == epochs - 1:
            model.eval()
            with torch.no_grad(), torch.autocast("cpu", dtype=torch.bfloat16):
                pv = model(X[va]).float()
            vt, ve = losses(torch.from_numpy(va), pv)
            hist.append((ep, round(tl.item(), 6), round(vt.item(), 6), round(ve.item(), 4)))
            print(f"[train] epoch {ep} tail {tl.item():.5f} val tail {vt.item():.5f} val estimate {ve.item():.4f} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    save(model, out, {"epochs": epochs, "samples": int(len(tr)), "val": int(nval), "history": hist})
    return model


def pack(model):
    """(bf16 weights as uint16, fp32 biases) in layer order conv1..conv4, fc1, fc2."""
    ws, bs = [], []
    for m in model:
        if hasattr(m, "weight"):
            ws.append(bf16_bits(m.weight.detach().float().contiguous().view(-1).numpy()))
            bs.append(m.bias.detach().float().numpy())
    return np.concatenate(ws), np.concatenate(bs).astype(np.float32)


def bf16_bits(w: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even FP32 -> BF16 bit patterns."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return r.astype(np.uint16)


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def save(model, path, meta):
    w, b = pack(model)
    np.savez(path, weights_bf16=w, biases=b, meta=json.dumps(meta))


def reference_forward(weights_bf16: np.ndarray, biases: np.ndarray, x: np.ndarray) -> np.ndarray:
    """FP32 forward of the packed (BF16-rounded) weights, [B,2,375] -> [B,1375]; the
    check for the HIP kernels."""
    w = bf16_to_f32(weights_bf16)
    wo = bo = 0
    h = x.astype(np.float32)
    cin = 2
    for layer in range(4):
        W = w[wo: wo + CH * cin * KW].reshape(CH, cin, KW); wo += CH * cin * KW
        B = biases[bo: bo + CH]; bo += CH
        L_out = LENS[layer + 1]
        hp = np.pad(h, ((0, 0), (0, 0), (KW // 2, KW // 2)))
        cols = np.stack([hp[:, :, 2 * t: 2 * t + KW] for t in range(L_out)], axis=2)   # [B, cin, L, KW]
        h = np.maximum(np.einsum("bclk,ock->bol", cols, W) + B[None, :, None], 0).astype(np.float32)
        cin = CH
    h = h.reshape(len(h), -1)
    W1 = w[wo: wo + HIDDEN * CH * LENS[-1]].reshape(HIDDEN, -1); wo += HIDDEN * CH * LENS[-1]
    b1 = biases[bo: bo + HIDDEN]; bo += HIDDEN
    h = np.maximum(h @ W1.T + b1, 0)
    W2 = w[wo: wo + N_OUT * HIDDEN].reshape(N_OUT, HIDDEN)
    b2 = biases[bo: bo + N_OUT]
    return (h @ W2.T + b2).astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["train"])
    ap.add_argument("data", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    train(a.data, a.out, epochs=a.epochs, threads=a.threads)


if __name__ == "__main__":
    main()
