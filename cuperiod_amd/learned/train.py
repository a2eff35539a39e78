"""Training of the proposer on synthetic survey curves labelled by the plain search.

Each training curve is drawn from the survey generator (``cuperiod_amd.workloads``) with
seeds disjoint from the benchmark's; its target is the range of coarse cells within one
refinement window (2,000 grid steps) of the best frequency the plain, full-grid BLS search
finds (200,000 frequencies). Features come from the same coarse search the proposer runs
at inference. Training minimises the mean cross-entropy between the softmax over the
4,096 cells and a uniform target on that range, with Adam on the HIP training kernels.

Run on a GPU::

    python -m cuperiod_amd.learned.train --curves 32768 --out weights.npz
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

#: Training seeds start here (the benchmark batches use 20261203 + 1..5).
TRAIN_SEED = 20261203 + 1000
#: Grid steps per refinement window (the proposer's default).
FINE_PER_WINDOW = 2000


def log(msg: str) -> None:
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def labelled_chunk(seed: int, n_curves: int, n_freq: int, settings: BLSSettings):
    """Features ([n, L, 2] device), target cell ranges, best frequencies and window
    geometry (f_lo, dlog, half-width) of one chunk of synthetic curves."""
    import torch

    from cuperiod_amd.workloads import survey_curves

    curves, grids, _ = survey_curves(n_curves, seed, n_freq)
    cs = engine.pack_curves(curves)
    prep = lp.prepare_grids(cs, grids, settings)
    pb = engine.pack_plans(cs, prep.full, settings)
    out, summary = engine.run_packed(pb)
    res = engine.finish(pb, out, summary)
    del out
    cout, csum = engine.run_packed(prep.coarse)
    x = lp.features(prep.coarse, cout, csum)
    torch.cuda.synchronize()
    best_f = 1.0 / res.best_period
    f_lo, dlog, half = lp.window_geometry(grids, FINE_PER_WINDOW)
    label = lp.label_ranges(best_f, f_lo, dlog, half)
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
        log(f"labelled {start + n}/{n_curves} curves ({time.perf_counter() - t0:.0f} s)")
    return (torch.cat(xs), np.concatenate(labels).astype(np.int32), np.concatenate(freqs),
            np.concatenate(geoms))


def window_recall(net: ProposerNet, x, freqs: np.ndarray, geom: np.ndarray,
                  max_windows: int = 32, radius: int = 1, batch: int = 512) -> dict:
    """Fraction of curves whose full-grid best frequency lies within a refinement window
    of one of the ``max_windows`` windows, and the mean probability the windows hold."""
    import torch

    hits, masses = [], []
    for s in range(0, x.shape[0], batch):
        g = geom[s:s + batch]
        gd = [torch.from_numpy(np.ascontiguousarray(g[:, j])).to(x.device) for j in range(3)]
        centre, mass = net.windows(x[s:s + batch].contiguous(), max_windows, radius, *gd)
        c = centre.cpu().numpy()
        m = mass.cpu().numpy()
        fc = g[:, 0:1] * np.exp(np.maximum(c, 0) * g[:, 1:2])
        near = np.abs(fc - freqs[s:s + batch, None]) <= g[:, 2:3]
        hits.append(np.any(near & (c >= 0), axis=1))
        masses.append(m.sum(axis=1))
    return {"recall": float(np.mean(np.concatenate(hits))),
            "mean_mass": float(np.mean(np.concatenate(masses)))}


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
        log(f"epoch {ep + 1}/{epochs}: loss {mean_loss:.3f}, val recall {rec['recall']:.3f}, "
            f"mass {rec['mean_mass']:.3f} ({time.perf_counter() - t0:.0f} s)")
    return net.parameters(), {"history": history, "n_train": int(tr.size), "n_val": int(n_val),
                              "steps": steps}


def main() -> int:
    ap = argparse.ArgumentParser(description="train the BLS candidate proposer")
    ap.add_argument("--curves", type=int, default=32768)
    ap.add_argument("--chunk", type=int, default=2048)
    ap.add_argument("--freqs", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=TRAIN_SEED)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    settings = BLSSettings(n_durations=6)
    x, labels, freqs, geom = build_set(args.curves, args.seed, args.chunk, args.freqs, settings)
    params, report = train(x, labels, freqs, geom, epochs=args.epochs, batch=args.batch,
                           lr=args.lr, seed=args.seed)
    save_parameters(params, args.out)
    report.update({"curves": args.curves, "freqs": args.freqs, "seed": args.seed,
                   "epochs": args.epochs, "batch": args.batch, "lr": args.lr})
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
