#!/usr/bin/env python3
"""Trains the learned preconditioner's conv net on the GPU at a given budget
and writes its parameters (.npy, FP32, the flat layout of numa_hip_net) and
the training record (.json) next to each other.

The pairs are generated by the solver itself (mixed Newton-FGMRES with the
V-cycle on random Bratu3D instances, lambda ~ U[1, 6.7], u0 = 0.1 U[0,1));
each (residual, V-cycle correction) pair feeds one Adam step of `batch`
pairs and is then dropped.  The budget is consumed in chunks of
`--chunk` pairs (each with its own instance seed), so progress is printed
between chunks; the step size decays linearly from --lr to --lr/10 over the
whole budget.

    python3 chebyshev-nonlinear-lab_amd/train_convnet.py --pairs 200000 \\
        --sizes 16,24,32 --out chebyshev-nonlinear-lab_amd/weights/convnet_200k_n16-32

Both committed parameter sets were written this way: 200k pairs on 16^3-32^3
grids (86 s on one MI355X; the default of the benchmark) and on 64^3-128^3
grids (596 s; a worse preconditioner: its 13^3 receptive field sees a smaller
part of the smooth error at those sizes).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--chunk", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--sizes", default="64,96,128")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--net-seed", type=int, default=21)
    ap.add_argument("--out", required=True, help="output path without extension")
    args = ap.parse_args()

    import numpy as np
    import numa_hip as H

    H.load()
    sizes = tuple(int(v) for v in args.sizes.split(","))
    ctx = H.Context(0)
    net = H.ConvNet(ctx, seed=args.net_seed)
    lr_end = args.lr / 10.0
    chunks = []
    done, t_all = 0, time.perf_counter()
    hist_all = []
    c = 0
    while done < args.pairs:
        p = min(args.chunk, args.pairs - done)
        lr0 = args.lr + (lr_end - args.lr) * done / args.pairs
        lr1 = args.lr + (lr_end - args.lr) * (done + p) / args.pairs
        cfg = H.net_train_config(pairs=p, batch=args.batch, sizes=sizes, lr=lr0, lr_end=lr1,
                                 seed=args.seed + 1000 * c, mg=H.mg_cheb(levels=0),
                                 loss=H.LOSS_RESIDUAL)
        st, hist = net.train_solver(cfg, hist_len=p // args.batch + 8)
        hist_all.append(hist)
        chunks.append({"pairs": int(st.pairs), "steps": int(st.steps), "solves": int(st.solves),
                       "seconds": round(st.seconds, 3), "loss_mean": float(hist.mean()),
                       "lr": [lr0, lr1]})
        done += int(st.pairs)
        c += 1
        print(f"chunk {c}: {done}/{args.pairs} pairs, {st.solves} solves, {st.seconds:.1f} s, "
              f"mean loss {hist.mean():.4f}", flush=True)
        save(args.out, net, np, chunks, hist_all, done, args, sizes, lr_end, t_all)
    rec = save(args.out, net, np, chunks, hist_all, done, args, sizes, lr_end, t_all)
    print(json.dumps({k2: v for k2, v in rec.items() if k2 != "chunks"}), flush=True)
    net.close()
    ctx.close()
    return 0


def save(out, net, np, chunks, hist_all, done, args, sizes, lr_end, t_all):
    """Parameters and the record so far (rewritten after every chunk)."""
    h = np.concatenate(hist_all)
    k = max(1, len(h) // 20)
    rec = {"pairs": done, "steps": int(len(h)), "batch": args.batch, "sizes": list(sizes),
           "lambda": [1.0, 6.7], "u0_noise": 0.1, "lr": [args.lr, lr_end], "loss": "residual",
           "seconds": round(time.perf_counter() - t_all, 1),
           "loss_first_window": float(h[:k].mean()), "loss_last_window": float(h[-k:].mean()),
           "solves": int(sum(ch["solves"] for ch in chunks)), "chunks": chunks,
           "net_seed": args.net_seed, "instance_seed": args.seed}
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.save(out + ".npy", net.get_params().astype(np.float32))
    with open(out + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    return rec


if __name__ == "__main__":
    sys.exit(main())
