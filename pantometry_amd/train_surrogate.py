"""Trains the heat surrogate on the GPU and writes its parameters.

    python -m pantometry_amd.train_surrogate [--samples 4096] [--edge 64] [--steps 1536]
                                             [--batch 32] [--lr 1e-3] [--rollout-pairs 0]
                                             [--rounds 1] [--unroll 3] [--init PATH]
                                             [--out PATH]

The training set is `samples` seeded trajectories of heterogeneous 64^3 blocks (lognormal(0, 0.5)
conductivity, eight hot spots), each advanced by the FP64 GPU stencil to a seeded multiple of 16
steps below 1024 and recorded across the next 16 steps; the network is trained with Adam (batch
32, lr 1e-3 on a half-cosine decay over all steps) by the library's own bf16 matrix-core kernels.
With --rollout-pairs N the steps are split into `rounds` + 1 stages; after each of the first
`rounds` stages N pairs are (re)built from the current network's own rollouts (1 to --unroll
surrogate steps from a stencil state, against the stencil from there) and trained on beside the
stencil pairs.  --init starts from saved parameters (Adam's moments start at zero again).

The shipped weights (pantometry_amd/weights/heat_unet_k16.npz) were made by four runs in turn,
about 24 minutes on one MI355X:

    python -m pantometry_amd.train_surrogate --steps 6000 --rollout-pairs 2048 --rounds 2 \
        --out stage1.npz
    python -m pantometry_amd.train_surrogate --init stage1.npz --lr 3e-4 --steps 8000 \
        --rollout-pairs 4096 --rounds 3 --out stage2.npz
    python -m pantometry_amd.train_surrogate --init stage2.npz --lr 2e-4 --steps 8000 \
        --rollout-pairs 4096 --rounds 3 --unroll 6 --out stage3.npz
    python -m pantometry_amd.train_surrogate --init stage3.npz --lr 1e-4 --steps 8000 \
        --rollout-pairs 4096 --rounds 3 --unroll 8
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
    from pantometry_amd.surrogate import (SURROGATE_SEED, WEIGHTS, Surrogate, add_rollout_pairs,
                                          build_training_set, train)
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--edge", type=int, default=64)
    ap.add_argument("--chunks", type=int, default=64, help="trajectory offsets, in 16-step chunks")
    ap.add_argument("--steps", type=int, default=1536)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rollout-pairs", type=int, default=0,
                    help="pairs built from the network's own rollouts (0: stencil pairs only)")
    ap.add_argument("--rounds", type=int, default=1, help="times the rollout pairs are rebuilt")
    ap.add_argument("--unroll", type=int, default=3,
                    help="most surrogate steps before a rollout pair is recorded")
    ap.add_argument("--seed", type=int, default=SURROGATE_SEED)
    ap.add_argument("--init", type=Path, default=None,
                    help="start from these parameters instead of the seeded initialisation")
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", flush=True)

    s = Surrogate(args.seed)
    if args.init is not None:
        s.load(args.init)
        log(f"starting from {args.init}")
    log(f"{s.parameter_count} parameters; building {args.samples} pairs at {args.edge}^3")
    build_s = build_training_set(s, args.samples, args.edge, args.chunks, log=log,
                                 capacity=args.samples + args.rollout_pairs)
    log(f"training set built in {build_s:.1f} s")
    t0 = time.perf_counter()
    stages = args.rounds + 1 if args.rollout_pairs else 1
    per = args.steps // stages
    losses, done, n = [], 0, args.samples
    for stage in range(stages):
        steps = per if stage < stages - 1 else args.steps - done
        if stage > 0:
            sec = add_rollout_pairs(s, args.samples, args.rollout_pairs, args.edge, args.chunks,
                                    unroll=args.unroll, log=log)
            n = args.samples + args.rollout_pairs
            log(f"round {stage}: {args.rollout_pairs} rollout pairs in {sec:.1f} s")
        losses += train(s, n, steps, args.batch, args.lr, log=log,
                        log_every=max(1, args.steps // 24), start=done, total=args.steps)
        done += steps
    train_s = time.perf_counter() - t0
    tail = sum(losses[-32:]) / min(32, len(losses))
    log(f"{args.steps} steps in {train_s:.1f} s; loss over the last 32 steps {tail:.3e}")
    s.save(args.out, samples=args.samples, edge=args.edge, chunks=args.chunks, steps=args.steps,
           batch=args.batch, lr=args.lr, seed=args.seed, final_loss=tail,
           rollout_pairs=args.rollout_pairs, rounds=args.rounds, unroll=args.unroll,
           init=str(args.init.name) if args.init is not None else "",
           build_seconds=round(build_s, 1), train_seconds=round(train_s, 1))
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
