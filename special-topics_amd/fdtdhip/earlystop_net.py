"""The early-stop network: definition, training and weight packing.

Architecture (fixed; the HIP inference kernels in csrc/earlystop_net.hpp implement
exactly this):
  input [2, 375]  (V and I prefix, mean-pooled by 4, each scaled by its max |x|)
  4 x Conv1d(k=5, stride=2, pad=2) + ReLU, 64 channels: lengths 375 -> 188 -> 94 -> 47 -> 24
  flatten 1536 -> Linear 512 + ReLU -> Linear 1375
  output: 2 x 687 pooled tail samples (V then I, each in units of the prefix's final
          amplitude: target = tail / end_amplitude) and log10 of the residual-energy ratio.
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


def load_data(paths):
    xs, ys = [], []
    for p in paths:
        d = np.load(p)
        x, tail, lr = d["x"], d["tail"], d["log_ratio"]
        a = end_amplitude(x)
        y = np.concatenate([(tail / a[:, :, None]).reshape(len(x), -1), lr[:, None]], axis=1)
        xs.append(x)
        ys.append(y)
    return np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32)


def train(paths, out, epochs=300, batch=64, lr=2e-3, val_frac=0.1, seed=0, threads=8):
    import torch
    torch.manual_seed(seed)
    torch.set_num_threads(threads)
    x, y = load_data(paths)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(x))
    nval = max(1, int(len(x) * val_frac))
    tr, va = perm[nval:], perm[:nval]
    X, Y = torch.from_numpy(x), torch.from_numpy(y)
    # Tail samples and the log-ratio scalar get comparable weight.
    wts = torch.ones(N_OUT)
    wts[-1] = float(2 * ES.N_OUT) * 0.05
    model = make_model()
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
            loss = (((pred.float() - Y[idx]) ** 2) * wts).mean()
            opt.zero_grad()
            loss.backward()
            opt.step()
            sched.step()
        if ep % 20 == 0 or ep == epochs - 1:
            model.eval()
            with torch.no_grad(), torch.autocast("cpu", dtype=torch.bfloat16):
                pv = model(X[va]).float()
            vl = float((((pv - Y[va]) ** 2) * wts).mean())
            hist.append((ep, round(float(loss), 6), round(vl, 6)))
            print(f"[train] epoch {ep} loss {float(loss):.5f} val {vl:.5f} ({time.time() - t0:.0f} s)", flush=True)
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
