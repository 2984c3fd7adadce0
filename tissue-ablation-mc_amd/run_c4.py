"""Full coupled ablation, chosen form vs plain (the 256^3 configuration).

The three-layer skin on a 256^3 grid over 2 x 2 x 1 cm runs 20 coupling
cycles of 5000 FP32 stencil steps each (FP64 Arrhenius inside every step,
ablation above 100 C and the optics/property rebuild per cycle):

  plain   1e8 packets per cycle;
  chosen  1e6 packets + the denoiser; a residual check compares 16^3-block
          sums of the denoised fluence with an independent 1e6-packet run
          (block sums on the device) and falls back to a 1e8-packet plain
          run for the cycle above 2%.

A third pass repeats the chosen form and, in every cycle, also runs 1e8
packets on the same tissue state (not timed) to score the denoised fluence.
Prints a JSON report: time per form, fallbacks, per-cycle fluence errors,
final peak temperatures, crater volumes.

    python tissue-ablation-mc_amd/run_c4.py [--cycles 20] [--steps 5000] [--n 256]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from tam_amd import configs, learned, lib  # noqa: E402
from tam_amd.denoise import Denoiser  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "denoiser_c3.f32")


def log(*a):
    print(*a, flush=True)


BLOCK = 16


def block_residual(A, B, frac=1e-2):
    """Median over blocks (B's sum above frac of the largest) of
    |A - B| / B, for block sums A (denoised) and B (independent run)."""
    sel = B > frac * B.max()
    return float(np.median(np.abs(A[sel] - B[sel]) / B[sel]))


def run(form, prob, args, dn=None, evaluate=False):
    p = prob.params
    g = prob.grid()
    g.tune_heat(reps=3)
    g.tune_mc(2_000_000)
    vcm = p.volume_voxel * 1e6
    rec = dict(form=form, cycles=[], fallbacks=0)
    g.synchronize()
    t_start = time.perf_counter()
    t_eval = 0.0                      # evaluation-only work and logging, not timed
    for c in range(args.cycles):
        cyc = dict(cycle=c)
        if form == "plain":
            st = g.mc_run(args.hi, seed=11, stream=c)
            npk = args.hi
        else:
            if evaluate:
                t0 = time.perf_counter()
                g.mc_run(args.hi, seed=11, stream=5000 + c)
                ref = g.download(lib.FIELD_TALLY) / (args.hi * vcm)
                t_eval += time.perf_counter() - t0
            g.mc_run(args.lo, seed=11, stream=3000 + c)
            B = g.tally_blocks(BLOCK)
            if evaluate:
                t0 = time.perf_counter()
                ind = g.download(lib.FIELD_TALLY)
                t_eval += time.perf_counter() - t0
            st = g.mc_run(args.lo, seed=11, stream=c)
            dn.apply(g, args.lo)
            A = g.tally_blocks(BLOCK)
            resid = block_residual(A, B)
            cyc["residual"] = resid
            npk = args.lo
            if evaluate:
                t0 = time.perf_counter()
                den = g.download(lib.FIELD_TALLY)
                phi_d, phi_i = den / (args.lo * vcm), ind / (args.lo * vcm)
                cyc["fluence_error_denoised"] = learned.fluence_error(phi_d, ref)
                cyc["fluence_error_plain_1e6"] = learned.fluence_error(phi_i, ref)
                t_eval += time.perf_counter() - t0
            if resid > args.tol:
                rec["fallbacks"] += 1
                st = g.mc_run(args.hi, seed=11, stream=c)
                npk = args.hi
                cyc["fallback"] = True
        g.normalise(prob.power / (npk * p.volume_voxel))
        if c in (1, 3, 7, 15):
            # the damage window fills over the first cycles and the fastest
            # stencil shape moves with it: re-time the shapes on the state
            g.tune_heat(reps=3)
        g.heat_run(args.steps)
        g.rebuild()
        g.synchronize()
        t0 = time.perf_counter()
        T = g.download(lib.FIELD_TEMP)[1:-1, 1:-1, 1:-1]
        m = g.download(lib.FIELD_MAT)
        tissue = m != len(prob.table)
        tm = g.timing()
        cyc.update(npk=npk, max_T_tissue=float(T[tissue].max()), removed=int((~tissue).sum()),
                   heat_ms=tm.heat_ms, heat_shape=[int(tm.tile), int(tm.kz)])
        rec["cycles"].append(cyc)
        log(form, json.dumps(cyc))
        t_eval += time.perf_counter() - t0
    g.synchronize()
    rec["wall_s"] = time.perf_counter() - t_start - t_eval
    rec["T"] = g.download(lib.FIELD_TEMP)[1:-1, 1:-1, 1:-1]
    rec["mat"] = g.download(lib.FIELD_MAT)
    g.close()
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--hi", type=int, default=100_000_000)
    ap.add_argument("--lo", type=int, default=1_000_000)
    ap.add_argument("--tol", type=float, default=0.02)
    ap.add_argument("--power", type=float, default=1.0)
    ap.add_argument("--report", default="c4_report.json")
    ap.add_argument("--no-eval", action="store_true")
    args = ap.parse_args()
    n = args.n
    prob = configs.config4_tissue(n=(n, n, n), packets=args.hi, steps=args.steps, cycles=args.cycles,
                                  power=args.power)
    dn = Denoiser((n, n, n), capacity=0)
    dn.set_params(np.fromfile(WEIGHTS, dtype=np.float32))
    plain = run("plain", prob, args)
    chosen = run("chosen", prob, args, dn=dn)
    ev = None if args.no_eval else run("chosen-eval", prob, args, dn=dn, evaluate=True)
    dn.close()
    vox_cm3 = (2.0 / n) * (2.0 / n) * (1.0 / n)
    air = len(prob.table)
    tp, tc = plain["mat"] != air, chosen["mat"] != air        # voxels still tissue
    out = dict(grid=n, cycles=args.cycles, steps=args.steps, hi=args.hi, lo=args.lo, tol=args.tol,
               plain_s=plain["wall_s"], chosen_s=chosen["wall_s"], speedup=plain["wall_s"] / chosen["wall_s"],
               fallbacks=chosen["fallbacks"],
               peak_T_plain=float(plain["T"][tp].max()), peak_T_chosen=float(chosen["T"][tc].max()),
               peak_T_error=float(abs(plain["T"][tp].max() - chosen["T"][tc].max())),
               max_abs_T_diff_tissue=float(np.abs(plain["T"] - chosen["T"])[tp & tc].max()),
               median_abs_T_diff_tissue=float(np.median(np.abs(plain["T"] - chosen["T"])[tp & tc])),
               crater_cm3_plain=float((plain["mat"] == air).sum() * vox_cm3),
               crater_cm3_chosen=float((chosen["mat"] == air).sum() * vox_cm3),
               crater_voxels_symdiff=int(((plain["mat"] == air) != (chosen["mat"] == air)).sum()),
               plain_cycles=plain["cycles"], chosen_cycles=chosen["cycles"])
    if ev is not None:
        e = [c["fluence_error_denoised"][0] for c in ev["cycles"]]
        e1 = [c["fluence_error_plain_1e6"][0] for c in ev["cycles"]]
        out.update(eval_cycles=ev["cycles"], median_fluence_error_denoised=float(np.median(e)),
                   median_fluence_error_plain_1e6=float(np.median(e1)))
    with open(args.report, "w") as f:
        json.dump(out, f, indent=1)
    log(json.dumps({k: v for k, v in out.items() if not k.endswith("_cycles")}))


if __name__ == "__main__":
    main()
