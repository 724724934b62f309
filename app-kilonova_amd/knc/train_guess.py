"""Trains the learned c2p guess network on the GPU and writes its weights.

Training set: (conserved per volume, pressure) pairs from snapshots of the
SRHD jet-in-cloud run (1024 x 256 zones, 8 blocks) taken every
`--snapshot-every` steps, plus seeded samples over Gamma in [1, 100], rho in
[1e-8, 1], p/rho in [1e-6, 10] (gamma 4/3 and 5/3), 10^7 pairs by default.
The network is trained by the library's own kernels (knc_mlp_train) and
written as an .npz of float32 arrays (params, norm, epoch_loss).

    python -m knc.train_guess --out app-kilonova_amd/knc/guess_net.npz
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import knc  # noqa: E402
from knc import mlp  # noqa: E402

DEFAULT_WEIGHTS = os.path.join(HERE, "guess_net.npz")


def harvest_jet_in_cloud(snapshots, every):
    s = knc.setups.jet_in_cloud()
    ctx = knc.context_from_setup(s, launch_mode=knc.BATCHED)
    h = knc.hydro_from_spec(s.hydro)
    cons, p = [], []
    for k in range(snapshots):
        for _ in range(every // 10):
            ctx.advance(10)
        c, q = mlp.harvest_pairs(ctx, h)
        cons.append(c)
        p.append(q)
        print(f"snapshot {k + 1}/{snapshots}: {c.shape[0]} pairs at step {(k + 1) * every}", flush=True)
    cons = np.concatenate(cons)
    return cons, np.concatenate(p), np.full(cons.shape[0], s.hydro.gamma_law_index)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--synthetic", type=int, default=7_900_000)
    ap.add_argument("--snapshots", type=int, default=8)
    ap.add_argument("--snapshot-every", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()

    t0 = time.time()
    hc, hp, hg = harvest_jet_in_cloud(args.snapshots, args.snapshot_every)
    sc, sp, sg = mlp.synthetic_pairs(args.synthetic, seed=args.seed)
    cons = np.concatenate([hc, sc])
    p = np.concatenate([hp, sp])
    gam = np.concatenate([hg, sg])
    print(f"training pairs: {cons.shape[0]} ({hc.shape[0]} harvested, {sc.shape[0]} seeded)", flush=True)
    net = mlp.GuessNetwork(seed=args.seed)
    losses = net.train(cons, p, gam, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    print("epoch loss:", " ".join(f"{x:.3e}" for x in losses), flush=True)
    params, norm = net.get()
    np.savez(args.out, params=params, norm=norm, epoch_loss=losses,
             pairs=np.array([hc.shape[0], sc.shape[0]], np.int64))
    print(f"wrote {args.out} in {time.time() - t0:.1f} s", flush=True)

    # held-out check: Newton iterations from the plain and the learned guess
    vc, vp, vg = mlp.synthetic_pairs(200_000, seed=args.seed + 1)
    _, guess = net.predict(vc, vg)
    it0, it1 = [], []
    for g in np.unique(vg):
        m = vg == g
        hyd = knc.relativistic_hydro(gamma_law_index=float(g))
        _, c0, i0 = knc.try_to_primitive(hyd, vc[m])
        _, c1, i1 = knc.try_to_primitive(hyd, vc[m], guess[m])
        ok = (c0 == 0) & (c1 == 0)
        it0.append(i0[ok])
        it1.append(i1[ok])
    it0, it1 = np.concatenate(it0), np.concatenate(it1)
    print(f"held-out mean Newton iterations: plain {it0.mean():.3f} learned {it1.mean():.3f}", flush=True)


def load(path=DEFAULT_WEIGHTS, device=0):
    """A GuessNetwork with the weights written by this script."""
    if not os.path.exists(path):
        raise RuntimeError(f"no trained weights at {path}: run python -m knc.train_guess")
    z = np.load(path, allow_pickle=False)
    net = mlp.GuessNetwork(device=device)
    net.set(z["params"], z["norm"])
    return net


if __name__ == "__main__":
    main()
