"""Trains the heat surrogate on the GPU and writes its parameters.

    python -m pantometry_amd.train_surrogate [--samples 4096] [--edge 64] [--steps 1536]
                                             [--batch 32] [--lr 1e-3] [--out PATH]

The training set is `samples` seeded trajectories of heterogeneous 64^3 blocks (lognormal(0, 0.5)
conductivity, eight hot spots), each advanced by the FP64 GPU stencil to a seeded multiple of 16
steps below 1024 and recorded across the next 16 steps; the network is trained with Adam (batch
32, lr 1e-3 on a half-cosine decay) by the library's own bf16 matrix-core kernels.  The defaults
are the published configuration; the run takes a few minutes on one MI355X.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def main(argv=None) -> int:
    from pantometry_amd.surrogate import (SURROGATE_SEED, WEIGHTS, Surrogate, build_training_set,
                                          train)
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--edge", type=int, default=64)
    ap.add_argument("--chunks", type=int, default=64, help="trajectory offsets, in 16-step chunks")
    ap.add_argument("--steps", type=int, default=1536)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=SURROGATE_SEED)
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", flush=True)

    s = Surrogate(args.seed)
    log(f"{s.parameter_count} parameters; building {args.samples} pairs at {args.edge}^3")
    build_s = build_training_set(s, args.samples, args.edge, args.chunks, log=log)
    log(f"training set built in {build_s:.1f} s")
    t0 = time.perf_counter()
    losses = train(s, args.samples, args.steps, args.batch, args.lr, log=log,
                   log_every=max(1, args.steps // 24))
    train_s = time.perf_counter() - t0
    tail = sum(losses[-32:]) / min(32, len(losses))
    log(f"{args.steps} steps in {train_s:.1f} s; loss over the last 32 steps {tail:.3e}")
    s.save(args.out, samples=args.samples, edge=args.edge, chunks=args.chunks, steps=args.steps,
           batch=args.batch, lr=args.lr, seed=args.seed, final_loss=tail,
           build_seconds=round(build_s, 1), train_seconds=round(train_s, 1))
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
