"""Train the fluence denoiser on MC-generated pairs and evaluate it.

    python tissue-ablation-mc_amd/train_denoiser.py --train 240 --eval 8 --steps 3000 \
        --out weights.f32 --report report.json

Training pairs: 128^3 layered tissues (tam_amd.learned), inputs from 1e5 or
2.5e5 packets, targets from --hi-train packets; evaluation pairs use
--hi-eval packets and are never trained on.  Writes the FP32 parameter vector
(tam_dn_get_params layout) and a JSON report: losses, evaluation L1 in log
space and per-voxel fluence error of the noisy input and of the prediction
against the evaluation targets, data-generation and training times.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from tam_amd import learned  # noqa: E402
from tam_amd.denoise import Denoiser  # noqa: E402


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--train", type=int, default=240)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--hi-train", type=float, default=1e7)
    ap.add_argument("--hi-eval", type=float, default=1e8)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="denoiser.f32")
    ap.add_argument("--report", default="denoiser_report.json")
    args = ap.parse_args()
    n = args.n
    dn = Denoiser((n, n, n), capacity=args.train + args.eval, seed=args.seed)
    t0 = time.perf_counter()
    meta = learned.make_pairs(dn, range(args.train), n, seed=1000 + args.seed, npk_hi=int(args.hi_train), log=log)
    t_gen_train = time.perf_counter() - t0
    t0 = time.perf_counter()
    meta_eval = learned.make_pairs(dn, range(args.train, args.train + args.eval), n, seed=5000 + args.seed,
                                   npk_hi=int(args.hi_eval), log=log)
    t_gen_eval = time.perf_counter() - t0
    before = [dn.eval(s) for s in range(args.train, args.train + args.eval)]
    # three phases of the learning rate (1, 1/3, 1/10), 100-step chunks so the
    # run reports progress
    losses = []
    t_train = 0.0
    phases = [(0.5, args.lr), (0.3, args.lr / 3), (0.2, args.lr / 10)]
    for frac, lr in phases:
        todo = int(round(frac * args.steps))
        while todo > 0:
            k = min(100, todo)
            t1 = time.perf_counter()
            l = dn.train(args.train, k, lr=lr, seed=args.seed * 7919 + len(losses))
            t_train += time.perf_counter() - t1
            losses.extend(float(x) for x in l)
            todo -= k
            log(f"step {len(losses)}: lr {lr:.1e}, mean loss of the last 100 {np.mean(losses[-100:]):.5f}, "
                f"{dn.last_ms() / k:.2f} ms/step")
    after = [dn.eval(s) for s in range(args.train, args.train + args.eval)]
    fwd_ms = []
    for s in range(args.train, args.train + args.eval):
        dn.eval(s)
        fwd_ms.append(dn.last_ms())
    # per-voxel fluence error against the evaluation targets, over voxels
    # above 1% of the peak: the noisy input and the prediction
    ferr = []
    for s in range(args.train, args.train + args.eval):
        pred, tgt, inp = dn.predict_slot(s)
        phi_t = np.expm1(tgt.astype(np.float64))
        ferr.append(dict(slot=s, pred=learned.fluence_error(np.expm1(pred.astype(np.float64)), phi_t),
                         input=learned.fluence_error(np.expm1(inp.astype(np.float64)), phi_t)))
        log(f"eval slot {s}: median/p90 fluence error input {ferr[-1]['input'][:2]} -> "
            f"denoised {ferr[-1]['pred'][:2]}")
    p = dn.params()
    p.tofile(args.out)
    flops = learned.conv_flops((n, n, n))
    step_ms = 1e3 * t_train / max(1, len(losses))
    rep = dict(grid=n, train_pairs=args.train, eval_pairs=args.eval, hi_train=args.hi_train,
               hi_eval=args.hi_eval, steps=len(losses), gen_train_s=t_gen_train, gen_eval_s=t_gen_eval,
               train_s=t_train, step_ms=step_ms, forward_ms=float(np.median(fwd_ms)),
               forward_tflops=flops / (np.median(fwd_ms) * 1e-3) / 1e12,
               train_tflops=3 * flops / (step_ms * 1e-3) / 1e12,
               loss_first100=float(np.mean(losses[:100])), loss_last100=float(np.mean(losses[-100:])),
               eval_l1_input=[b[1] for b in before], eval_l1_before=[b[0] for b in before],
               eval_l1_after=[a[0] for a in after], fluence_error=ferr,
               median_fluence_error_input=float(np.median([e["input"][0] for e in ferr])),
               median_fluence_error_pred=float(np.median([e["pred"][0] for e in ferr])),
               pairs=meta, eval_meta=meta_eval)
    with open(args.report, "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k not in ("pairs", "eval_meta")}))
    dn.close()


if __name__ == "__main__":
    main()
