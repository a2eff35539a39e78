"""Training of the proposer on synthetic gapped curves labelled by the plain search.

Each training curve is drawn from the gapped-curve recipe (``workloads.gapped_curves``:
N = 20,000 over 1,000 d less 30% in seasonal gaps, sigma 2e-3, a 0.2 d transit of depth
1.5e-3 at a period log-uniform in [2, 20] d) with seeds disjoint from the benchmark's;
its target is the coarse cell nearest the best frequency the plain, full-grid BLS search
finds (200,000 frequencies), +-1 cell.
Features come from the same chunked coarse search the proposer runs at inference.
Training minimises the mean cross-entropy between the softmax over the 4,096 cells and
a uniform target on those three cells, with Adam on the HIP training kernels.

Run on a GPU::

    python -m cuperiod_amd.learned.train --curves 16384 --out weights.npz
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned import proposer as lp
from cuperiod_amd.learned.network import ProposerNet, init_parameters, save_parameters

#: Training seeds start here (the benchmark batches use 20261203 + 1..8).
TRAIN_SEED = 20261203 + 1000
#: Grid steps per refinement window (the proposer's default).
FINE_PER_WINDOW = 2000


def log(msg: str) -> None:
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def labelled_chunk(seed: int, n_curves: int, n_freq: int, settings: BLSSettings):
    """Features ([n, L, 2] device), target cell ranges, best frequencies and window
    geometry (f_lo, dlog, half-width) of one chunk of synthetic curves."""
    import torch

    from cuperiod_amd.workloads import gapped_curves

    curves, grids, _ = gapped_curves(n_curves, seed, n_freq)
    cs = engine.pack_curves(curves)
    prep = lp.prepare_grids(cs, grids, settings, fine_per_window=FINE_PER_WINDOW)
    out, summary = engine.run_packed(prep.full)
    res = engine.finish(prep.full, out, summary)
    del out
    x = lp.coarse_features(prep)
    torch.cuda.synchronize()
    best_f = 1.0 / res.best_period
    f_lo, dlog, half = prep.geometry_host
    label = lp.label_cells(best_f, f_lo, dlog, 1)
    geom = np.stack([f_lo, dlog, half], axis=1)
    return x, label, best_f, geom


def build_set(n_curves: int, seed: int, chunk: int, n_freq: int, settings: BLSSettings):
    import torch

    xs, labels, freqs, geoms = [], [], [], []
    t0 = time.perf_counter()
    for i, start in enumerate(range(0, n_curves, chunk)):
        n = min(chunk, n_curves - start)
        x, lab, f, g = labelled_chunk(seed + i, n, n_freq, settings)
        xs.append(x)
        labels.append(lab)
        freqs.append(f)
        geoms.append(g)
        torch.cuda.empty_cache()
        log(f"labelled {start + n}/{n_curves} curves ({time.perf_counter() - t0:.0f} s)")
    return (torch.cat(xs), np.concatenate(labels).astype(np.int32), np.concatenate(freqs),
            np.concatenate(geoms))


def window_recall(net: ProposerNet, x, freqs: np.ndarray, geom: np.ndarray,
                  max_windows: int = 32, radius: int = 1, batch: int = 512,
                  mass: float = 0.995) -> dict:
    """Recall of the full-grid best frequency: within the refined range of any of the
    ``max_windows`` windows (every frequency within the half-width of a claimed cell),
    and within the windows kept until their probability reaches ``mass``."""
    import torch

    hits, kept_hits, masses, kept_counts = [], [], [], []
    for s in range(0, x.shape[0], batch):
        g = geom[s:s + batch]
        # windows claim centre +-radius cells; refinement adds the half-width g[:, 2]
        gd = [torch.from_numpy(np.ascontiguousarray(g[:, j] if j < 2 else 0.0 * g[:, j]))
              .to(x.device) for j in range(3)]
        centre, m, span = net.windows(x[s:s + batch].contiguous(), max_windows, radius, *gd)
        c = centre.cpu().numpy()
        m = np.where(c >= 0, m.cpu().numpy(), 0.0)
        sp = span.cpu().numpy()
        f_a = g[:, 0:1] * np.exp(np.maximum(sp[..., 0], 0) * g[:, 1:2]) - g[:, 2:3]
        f_b = g[:, 0:1] * np.exp(np.maximum(sp[..., 1], 0) * g[:, 1:2]) + g[:, 2:3]
        f = freqs[s:s + batch, None]
        inside = (f >= f_a) & (f <= f_b) & (c >= 0)
        order = np.argsort(-m, axis=1, kind="stable")
        ms = np.take_along_axis(m, order, axis=1)
        before = np.concatenate([np.zeros((m.shape[0], 1)), np.cumsum(ms, axis=1)[:, :-1]], axis=1)
        keep = np.zeros_like(inside)
        np.put_along_axis(keep, order, (before < mass) & (ms > 0), axis=1)
        hits.append(inside.any(axis=1))
        kept_hits.append((inside & keep).any(axis=1))
        masses.append(np.where(keep, m, 0.0).sum(axis=1))
        kept_counts.append(keep.sum(axis=1))
    return {"recall": float(np.mean(np.concatenate(hits))),
            "kept_recall": float(np.mean(np.concatenate(kept_hits))),
            "mean_kept_mass": float(np.mean(np.concatenate(masses))),
            "mean_kept_windows": float(np.mean(np.concatenate(kept_counts)))}


def train(x, labels: np.ndarray, freqs: np.ndarray, geom: np.ndarray, *, epochs: int,
          batch: int, lr: float, seed: int, val_frac: float = 0.1) -> tuple[dict, dict]:
    import torch

    n = x.shape[0]
    n_val = max(1, int(n * val_frac))
    idx = np.random.default_rng(seed).permutation(n)
    val, tr = idx[:n_val], idx[n_val:]
    net = ProposerNet(init_parameters(seed))
    lab_d = torch.from_numpy(labels).to(x.device)
    rng = np.random.default_rng(seed + 1)
    steps = 0
    total = epochs * (tr.size // batch)
    history = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(tr)
        net.loss.zero_()
        for s in range(0, order.size - batch + 1, batch):
            sel = torch.from_numpy(order[s:s + batch]).to(x.device)
            xb = x.index_select(0, sel).contiguous()
            lb = lab_d.index_select(0, sel).contiguous()
            # cosine decay from lr to lr/20
            rate = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * steps / max(1, total))))
            net.train_step(xb, lb, rate)
            steps += 1
        torch.cuda.synchronize()
        mean_loss = net.loss.item() / max(1, order.size // batch)
        vsel = torch.from_numpy(val).to(x.device)
        rec = window_recall(net, x.index_select(0, vsel).contiguous(), freqs[val], geom[val])
        history.append({"epoch": ep + 1, "train_loss": mean_loss, **rec,
                        "seconds": time.perf_counter() - t0})
        log(f"epoch {ep + 1}/{epochs}: loss {mean_loss:.3f}, val recall {rec['recall']:.4f}, "
            f"kept {rec['kept_recall']:.4f} with {rec['mean_kept_windows']:.1f} windows "
            f"({time.perf_counter() - t0:.0f} s)")
    return net.parameters(), {"history": history, "n_train": int(tr.size), "n_val": int(n_val),
                              "steps": steps}


def main() -> int:
    ap = argparse.ArgumentParser(description="train the BLS candidate proposer")
    ap.add_argument("--curves", type=int, default=16384)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--freqs", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=TRAIN_SEED)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    settings = BLSSettings(n_durations=6)
    t0 = time.perf_counter()
    x, labels, freqs, geom = build_set(args.curves, args.seed, args.chunk, args.freqs, settings)
    label_s = time.perf_counter() - t0
    # how often the stacked coarse power alone peaks on the target (no network)
    top = x[:, :, 0].argmax(dim=1).cpu().numpy()
    centre = labels.mean(axis=1)
    feature_hit = float(np.mean(np.abs(top - centre) <= 1.0))
    log(f"stacked-feature argmax on the target cell: {feature_hit:.4f}")
    params, report = train(x, labels, freqs, geom, epochs=args.epochs, batch=args.batch,
                           lr=args.lr, seed=args.seed)
    report["feature_argmax_on_target"] = feature_hit
    save_parameters(params, args.out)
    report.update({"curves": args.curves, "freqs": args.freqs, "seed": args.seed,
                   "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
                   "labelling_seconds": round(label_s, 1),
                   "workload": "gapped curves (workloads.gapped_curves), plain-search labels"})
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
