"""The early-stop network: definition, training and weight packing.

Architecture (fixed; the HIP inference kernels in csrc/earlystop_net.hpp implement
exactly this):
  input [2, 375]  (V and I prefix, mean-pooled by 4, each scaled by its max |x|)
  4 x Conv1d(k=5, stride=2, pad=2) + ReLU, 64 channels: lengths 375 -> 188 -> 94 -> 47 -> 24
  flatten 1536 -> Linear 512 + ReLU -> Linear 1376
  output: 2 x 687 pooled tail samples (V then I, each in units of the prefix's final
          amplitude: target = tail / end_amplitude), an estimate of log10 of the
          residual energy -- the energy of the difference between the predicted and the
          true tail -- over the run's total port energy (prefix plus tail, in the
          prefix-normalised units of the input), and an estimate of log10 of the
          early-stopped S11's maximum error in dB over the 1000 frequencies (against the
          full run).  The fallback to the full-length FP64 run keys on the S11-error
          estimate, with a threshold calibrated on the validation split (calibrate()).
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
N_OUT = 2 * ES.N_OUT + 2
ERR_FLOOR_DB = 1e-3       # the S11-error target is log10(max |dB error| + ERR_FLOOR_DB)
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
    """x [n,2,375], y [n,1374] (scaled tail), a [n,2] end amplitudes, e_total [n], and
    for the S11-error target: channel scales [n,2], prefix DFTs V, I [n,F] and the full
    run's S11 [n,F] (complex), dt [n]."""
    xs, ys, amps, etot, scales, vps, ips, s11s, dts = [], [], [], [], [], [], [], [], []
    freqs = ES.frequency_grid({"maxFreq": 6e9, "step": 6e6})
    for p in paths:
        d = np.load(p)
        x, tail = d["x"], d["tail"]
        a = end_amplitude(x)
        ep, et = energies(x, tail)
        xs.append(x)
        ys.append((tail / a[:, :, None]).reshape(len(x), -1))
        amps.append(a)
        etot.append(ep + et)
        scales.append(d["scale"])
        dts.append(d["dt"])
        s11s.append(d["s11"])
        # The prefix DFTs of all variants with one dt as two matrix products.
        for dt in np.unique(d["dt"]):
            sel = np.nonzero(d["dt"] == dt)[0]
            w, n = 2 * np.pi * freqs, np.arange(ES.N_PREFIX)
            ev = float(dt) * np.exp(-1j * np.outer(w, (n + 1) * dt))
            ei = float(dt) * np.exp(-1j * np.outer(w, (n + 1) * dt - dt / 2))
            V = (ev @ d["v_prefix"][sel].astype(np.float64).T).T
            I = (ei @ d["i_prefix"][sel].astype(np.float64).T).T
            for q, k in enumerate(sel):
                vps.append((len(xs) - 1, k, V[q]))
                ips.append(I[q])
    # Back to the variants' order (file, then index within it).
    order = sorted(range(len(vps)), key=lambda q: vps[q][:2])
    vp = np.array([vps[q][2] for q in order])
    ip = np.array([ips[q] for q in order])
    return (np.concatenate(xs).astype(np.float32), np.concatenate(ys).astype(np.float32),
            np.concatenate(amps).astype(np.float32), np.concatenate(etot).astype(np.float32),
            {"scale": np.concatenate(scales), "vp": vp, "ip": ip,
             "s11": np.concatenate(s11s).astype(np.complex128), "dt": np.concatenate(dts), "freqs": freqs})


def s11_error_db(tail_scaled, a, aux, idx):
    """Max over frequency of |20 log10 |S11_early| - 20 log10 |S11_full||, where the
    early S11 is the prefix DFT plus the DFT of the pooled tail `tail_scaled`
    ([n, 1374] in end-amplitude units, numpy) of variants idx."""
    tails = tail_scaled.reshape(-1, 2, ES.N_OUT) * a[:, :, None]
    err = np.empty(len(idx))
    for q, k in enumerate(idx):
        Vt, It = ES.tail_dft(tails[q], aux["scale"][k], float(aux["dt"][k]), aux["freqs"])
        S = ES.s11_from(aux["vp"][k] + Vt, aux["ip"][k] + It)
        err[q] = np.max(np.abs(20 * np.log10(np.abs(S)) - 20 * np.log10(np.abs(aux["s11"][k]))))
    return err


def residual_log_ratio(pred_tail, true_tail, a, e_total):
    """log10 of the pooled-tail error energy over the total energy (torch tensors)."""
    import torch
    err = ((pred_tail - true_tail).view(-1, 2, ES.N_OUT) * a[:, :, None]) ** 2
    return torch.log10(ES.POOL_OUT * err.sum(dim=(1, 2)) / e_total + 1e-30)


def train(paths, out, epochs=300, batch=64, lr=2e-3, val_frac=0.2, seed=0, threads=8, target_within=0.95):
    """Train the tail and residual-energy outputs on (1 - val_frac) of the variants;
    then fit the S11-error output and calibrate its fallback threshold on the rest
    (fit_error_head, calibrate), and save the BF16-packed weights with the threshold."""
    import torch
    torch.manual_seed(seed)
    torch.set_num_threads(threads)
    x, y, a, et, aux = load_data(paths)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(x))
    nval = max(2, int(len(x) * val_frac))
    tr, va = perm[nval:], perm[:nval]
    X, Y, A, ET = torch.from_numpy(x), torch.from_numpy(y), torch.from_numpy(a), torch.from_numpy(et)
    n_tail = 2 * ES.N_OUT
    model = make_model()

    def losses(idx, pred):
        pt = pred[:, :n_tail].float()
        tail_loss = ((pt - Y[idx]) ** 2).mean()
        # Output n_tail learns the (detached) residual energy of the current prediction.
        target = residual_log_ratio(pt.detach(), Y[idx], A[idx], ET[idx])
        est_loss = ((pred[:, n_tail].float() - target) ** 2).mean()
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
    fit_idx, cal_idx = va[: nval // 2], va[nval // 2:]
    fit_error_head(model, x, a, aux, fit_idx)
    w, b = pack(model)
    cal = calibrate(w, b, x, a, aux, cal_idx, target_within)
    print(f"[train] calibration: {json.dumps(cal)}", flush=True)
    np.savez(out, weights_bf16=w, biases=b,
             meta=json.dumps({"epochs": epochs, "samples": int(len(tr)), "error_fit": int(len(fit_idx)),
                              "calibration": cal, "history": hist}))
    return model


def hidden(model, x):
    """The 512 ReLU activations before the last layer, FP32, numpy."""
    import torch
    with torch.no_grad():
        return model[:-1](torch.from_numpy(x)).float().numpy()


def fit_error_head(model, x, a, aux, idx, ridge=1e-2):
    """Set the last output (log10 of the S11 error in dB) by ridge regression on the
    frozen hidden layer over variants idx, whose true S11 errors come from the tail the
    network predicts for them."""
    import torch
    n_tail = 2 * ES.N_OUT
    with torch.no_grad():
        pred = model(torch.from_numpy(x[idx])).float().numpy()
    err = s11_error_db(pred[:, :n_tail], a[idx], aux, idx)
    t = np.log10(err + ERR_FLOOR_DB)
    H = hidden(model, x[idx]).astype(np.float64)
    Hm, tm = H.mean(0), t.mean()
    Hc = H - Hm
    wv = np.linalg.solve(Hc.T @ Hc + ridge * len(idx) * np.eye(H.shape[1]), Hc.T @ (t - tm))
    last = model[-1]
    with torch.no_grad():
        last.weight[-1].copy_(torch.from_numpy(wv.astype(np.float32)))
        last.bias[-1] = float(tm - Hm @ wv)


def calibrate(w, b, x, a, aux, idx, target_within=0.95):
    """Fallback threshold on the S11-error estimate of the packed (BF16) network over
    variants idx: the largest threshold whose non-fallback variants, with fallbacks
    counted as exact, keep at least target_within of them within 0.1 dB; plus the
    accuracy / fallback-rate curve."""
    n_tail = 2 * ES.N_OUT
    y = reference_forward(w, b, x[idx])
    err = s11_error_db(y[:, :n_tail], a[idx], aux, idx)
    est = y[:, -1]
    ok = err <= 0.1
    cands = np.concatenate([[-np.inf], np.sort(est)])
    best = -np.inf
    curve = []
    for th in cands:
        keep = est <= th
        within = float(np.mean(ok | ~keep))
        curve.append((float(th), float(np.mean(~keep)), within))
        if within >= target_within:
            best = th
    keep = est <= best
    pts = [curve[int(q)] for q in np.linspace(0, len(curve) - 1, 11)]
    return {"threshold": float(best), "variants": int(len(idx)), "fallback_rate": float(np.mean(~keep)),
            "within_0p1db": float(np.mean(ok | ~keep)), "within_0p1db_network_only": float(np.mean(ok)),
            "err_db_median": float(np.median(err)), "est_vs_err_corr": float(np.corrcoef(est, np.log10(err + ERR_FLOOR_DB))[0, 1]),
            "curve_threshold_fallback_within": pts}


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


def reference_forward(weights_bf16: np.ndarray, biases: np.ndarray, x: np.ndarray) -> np.ndarray:
    """FP32 forward of the packed (BF16-rounded) weights, [B,2,375] -> [B,1376]; the
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
