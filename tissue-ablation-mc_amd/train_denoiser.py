"""Train the fluence denoiser on MC-generated pairs and evaluate it.

    python tissue-ablation-mc_amd/train_denoiser.py \\
        --group layers:1300:1e7:1e5,2.5e5,1e6 --group skin:500:1e7:2.5e5,1e6 \\
        --eval-group layers:16:1e8:1e5 --steps 20000 --out weights.f32 --report report.json

A group is kind:count:target_packets:input_packets[,input_packets...]; kind
is "layers" (128^3 tissues of 1-4 random planar layers) or "skin" (jittered
three-layer skin), see tam_amd.learned.  Each group draws its tissues from its
own seed (training group i: 1000 + seed + 100 i, evaluation group j: 5000 +
seed + 100 j), so an evaluation group holds the same tissues whatever the
training set is.  Evaluation pairs are never trained on.

--init starts from a stored parameter vector; --steps 0 with --init only
evaluates.  --compare scores a second parameter vector on the same
evaluation pairs.  Writes the FP32 parameter vector (tam_dn_get_params
layout) and a JSON report: losses, evaluation L1 in log space and per-voxel
fluence error of the noisy input and of the prediction against the
evaluation targets, per evaluation group; data-generation and training times.
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


def parse_group(text):
    kind, count, hi, lo = text.split(":")
    if kind not in ("layers", "skin"):
        raise SystemExit(f"unknown tissue kind {kind!r} in group {text!r}")
    return dict(kind=kind, count=int(count), hi=int(float(hi)), lo=tuple(int(float(x)) for x in lo.split(",")))


def score(dn, slots):
    """Per-slot per-voxel fluence error (median, p90, voxels) of the input
    and of the prediction against the target, voxels above 1% of peak."""
    out = []
    for s in slots:
        pred, tgt, inp = dn.predict_slot(s)
        phi_t = np.expm1(tgt.astype(np.float64))
        phi_p = np.expm1(pred.astype(np.float64))
        out.append(dict(slot=s, pred=learned.fluence_error(phi_p, phi_t),
                        input=learned.fluence_error(np.expm1(inp.astype(np.float64)), phi_t),
                        bias=learned.fluence_bias(phi_p, phi_t)))
    return out


def summary(errs):
    pred = [e["pred"][0] for e in errs]
    inp = [e["input"][0] for e in errs]
    return dict(pairs=len(errs), median_fluence_error_input=float(np.median(inp)),
                median_fluence_error_pred=float(np.median(pred)),
                p90_over_pairs_fluence_error_pred=float(np.percentile(pred, 90)),
                worst_pair_fluence_error_pred=float(np.max(pred)),
                median_signed_bias_pred=float(np.median([e["bias"][0] for e in errs])),
                median_sum_bias_pred=float(np.median([e["bias"][1] for e in errs])),
                pairs_pred_better=int(sum(p < i for p, i in zip(pred, inp))))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--group", action="append", default=[], help="training group")
    ap.add_argument("--eval-group", action="append", default=[], help="evaluation group")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="starting parameter vector")
    ap.add_argument("--compare", default=None, help="second parameter vector scored on the evaluation pairs")
    ap.add_argument("--out", default="denoiser.f32")
    ap.add_argument("--report", default="denoiser_report.json")
    args = ap.parse_args()
    n = args.n
    train_groups = [parse_group(g) for g in args.group]
    eval_groups = [parse_group(g) for g in args.eval_group]
    ntrain = sum(g["count"] for g in train_groups)
    neval = sum(g["count"] for g in eval_groups)
    if args.steps > 0 and ntrain == 0:
        raise SystemExit("training steps without training pairs")
    dn = Denoiser((n, n, n), capacity=ntrain + neval, seed=args.seed)
    if args.init:
        dn.set_params(np.fromfile(args.init, dtype=np.float32))
    meta, meta_eval, gen_s = [], [], []
    slot = 0
    for i, g in enumerate(train_groups + eval_groups):
        is_eval = i >= len(train_groups)
        gseed = (5000 + 100 * (i - len(train_groups)) if is_eval else 1000 + 100 * i) + args.seed
        t0 = time.perf_counter()
        m = learned.make_pairs(dn, range(slot, slot + g["count"]), n, seed=gseed, npk_hi=g["hi"],
                               npk_lo_choices=g["lo"], log=log, kind=g["kind"])
        gen_s.append(dict(group=i, eval=is_eval, seed=gseed, seconds=time.perf_counter() - t0, **g))
        (meta_eval if is_eval else meta).append(m)
        g["slots"] = list(range(slot, slot + g["count"]))
        slot += g["count"]
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
            l = dn.train(ntrain, k, lr=lr, seed=args.seed * 7919 + len(losses))
            t_train += time.perf_counter() - t1
            losses.extend(float(x) for x in l)
            todo -= k
            log(f"step {len(losses)}: lr {lr:.1e}, mean loss of the last 100 {np.mean(losses[-100:]):.5f}, "
                f"{dn.last_ms() / k:.2f} ms/step")
    p = dn.params()
    p.tofile(args.out)
    evals = []
    for g in eval_groups:
        errs = score(dn, g["slots"])
        ev = dict(kind=g["kind"], hi=g["hi"], lo=g["lo"], **summary(errs), per_pair=errs)
        log(f"eval {g['kind']} x{g['count']} (inputs {g['lo']}): median fluence error input "
            f"{ev['median_fluence_error_input']:.4f} -> denoised {ev['median_fluence_error_pred']:.4f}")
        evals.append(ev)
    compared = []
    if args.compare:
        dn.set_params(np.fromfile(args.compare, dtype=np.float32))
        for g in eval_groups:
            ev = dict(kind=g["kind"], weights=args.compare, **summary(score(dn, g["slots"])))
            log(f"eval {g['kind']} with {args.compare}: denoised {ev['median_fluence_error_pred']:.4f}")
            compared.append(ev)
        dn.set_params(p)
    fwd_ms = []
    for s in range(ntrain, ntrain + min(neval, 8)):
        dn.eval(s)
        fwd_ms.append(dn.last_ms())
    flops = learned.conv_flops((n, n, n))
    step_ms = 1e3 * t_train / max(1, len(losses))
    rep = dict(grid=n, train_groups=[{k: v for k, v in g.items() if k != "slots"} for g in train_groups],
               eval_groups=[{k: v for k, v in g.items() if k != "slots"} for g in eval_groups],
               init=args.init, steps=len(losses), generation=gen_s, train_s=t_train, step_ms=step_ms,
               forward_ms=float(np.median(fwd_ms)) if fwd_ms else None,
               forward_tflops=flops / (np.median(fwd_ms) * 1e-3) / 1e12 if fwd_ms else None,
               train_tflops=3 * flops / (step_ms * 1e-3) / 1e12 if losses else None,
               loss_first100=float(np.mean(losses[:100])) if losses else None,
               loss_last100=float(np.mean(losses[-100:])) if losses else None,
               evaluation=evals, compared=compared, pairs=meta, eval_meta=meta_eval)
    with open(args.report, "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k not in ("pairs", "eval_meta", "evaluation")}))
    log(json.dumps([{k: v for k, v in e.items() if k != "per_pair"} for e in evals]))
    dn.close()


if __name__ == "__main__":
    main()
