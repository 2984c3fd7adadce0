"""Time the fluence denoiser: forward passes and training steps.

    python tissue-ablation-mc_amd/time_denoiser.py --n 128 --train-steps 200
    python tissue-ablation-mc_amd/time_denoiser.py --n 256 --train-steps 0

Forward passes run on random feature maps with the committed weights; the
training steps run on `--slots` stored pairs of random features (the target
is a scaled copy of the first feature).  Times come from the network's own
device timer (tam_dn_last_ms).
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from tam_amd.denoise import Denoiser  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "denoiser_c3.f32")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128, help="grid edge (a multiple of 32)")
    ap.add_argument("--forward", type=int, default=5, help="forward passes")
    ap.add_argument("--train-steps", type=int, default=200)
    ap.add_argument("--slots", type=int, default=8)
    args = ap.parse_args()
    n = args.n
    slots = args.slots if args.train_steps > 0 else 0
    dn = Denoiser((n, n, n), capacity=slots, seed=1)
    dn.set_params(np.fromfile(WEIGHTS, dtype=np.float32))
    rng = np.random.default_rng(0)
    for s in range(slots):
        f = rng.random((4, n, n, n)).astype(np.float32)
        dn.store_host(s, f, 1.01 * f[0])
    f = rng.random((4, n, n, n)).astype(np.float32)
    flop = 2.65e11 * (n / 128) ** 3
    for i in range(args.forward):
        dn.forward(f)
        ms = dn.last_ms()
        print(f"forward {n}^3: {ms:.3f} ms, {flop / (ms * 1e-3) / 1e12:.1f} TFLOP/s", flush=True)
    if args.train_steps > 0:
        dn.train(slots, args.train_steps, lr=1e-4, seed=0)
        ms = dn.last_ms() / args.train_steps
        print(f"training {n}^3: {ms:.3f} ms/step, {3 * flop / (ms * 1e-3) / 1e12:.1f} TFLOP/s", flush=True)
    dn.close()


if __name__ == "__main__":
    main()
