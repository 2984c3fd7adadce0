"""Batched uneven instances (the 64-tissue configuration): batched vs one at a time.

The committed manifest (tam_amd/batch_manifest.json): 64 homogeneous 1 cm
cubes, grid edges from {64, 96, 128, 160}, optics and packet counts drawn
once.  Each instance: its packets, normalisation, 2000 FP32 heat steps with
the config-0 thermal constants.

  batched     tam_batch_tune (transport shape, stencil shape), then one
              transport launch for all 64 instances and 2000 batched stencil
              launches;
  one-by-one  the same work with tam_mc_run / tam_normalise / tam_heat_run per
              instance, each with its own tuned launch shapes.

Prints a JSON report (packets/s, voxel-updates/s, wall and kernel times) and
checks that both forms give the same integer fates per instance.

    python tissue-ablation-mc_amd/run_c2.py [--steps 2000]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from tam_amd import configs, lib  # noqa: E402


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--report", default="c2_report.json")
    args = ap.parse_args()
    man = json.load(open(os.path.join(os.path.dirname(configs.__file__), "batch_manifest.json")))
    probs = configs.batch_problems(man, steps=args.steps)
    npk = [m["packets"] for m in man]
    vox = sum(p.params.nx * p.params.ny * p.params.nz for p in probs)
    log(f"{len(probs)} instances, {vox / 1e6:.1f} M voxels, {sum(npk) / 1e6:.1f} M packets")

    # ---- batched
    gs = [p.grid() for p in probs]
    b = lib.Batch(gs)
    t0 = time.perf_counter()
    shape = b.tune([max(20_000, n // 50) for n in npk], heat_reps=3)
    t_tune = time.perf_counter() - t0
    log("batch tuned", shape, f"{t_tune:.1f} s")
    t0 = time.perf_counter()
    st_b = b.mc_run(npk, seed=7, stream0=0)
    mc_ms_b = b.last_ms()
    for p, g in zip(probs, gs):
        g.normalise(p.scale)
    b.heat_run(args.steps)
    heat_ms_b = b.last_ms()
    for g in gs:
        g.synchronize()
    wall_b = time.perf_counter() - t0
    log(f"batched: transport {mc_ms_b:.1f} ms, heat {heat_ms_b:.1f} ms, wall {wall_b:.2f} s")
    Tb = [g.download(lib.FIELD_TEMP) for g in gs[:4]]
    b.close()
    for g in gs:
        g.close()

    # ---- one at a time (each instance tuned on its own)
    gs = [p.grid() for p in probs]
    t0 = time.perf_counter()
    for g, n in zip(gs, npk):
        g.tune_mc(max(20_000, n // 50))
        g.tune_heat(reps=3)
    t_tune1 = time.perf_counter() - t0
    mc_ms_1 = heat_ms_1 = 0.0
    st_1 = []
    t0 = time.perf_counter()
    for i, (p, g) in enumerate(zip(probs, gs)):
        st_1.append(g.mc_run(npk[i], seed=7, stream=i))
        mc_ms_1 += st_1[-1].kernel_ms
        g.normalise(p.scale)
        g.heat_run(args.steps)
        heat_ms_1 += g.timing().heat_ms
    for g in gs:
        g.synchronize()
    wall_1 = time.perf_counter() - t0
    log(f"one-by-one: transport {mc_ms_1:.1f} ms, heat {heat_ms_1:.1f} ms, wall {wall_1:.2f} s")
    same = all(a.absorbed == c.absorbed and a.crossings == c.crossings and a.interactions == c.interactions
               for a, c in zip(st_b, st_1))
    dT = max(float(np.abs(t - g.download(lib.FIELD_TEMP)).max()) for t, g in zip(Tb, gs[:4]))
    for g in gs:
        g.close()
    out = dict(instances=len(probs), voxels=vox, packets=sum(npk), steps=args.steps, batch_shape=shape,
               batch_tune_s=t_tune, single_tune_s=t_tune1,
               batched=dict(mc_ms=mc_ms_b, heat_ms=heat_ms_b, wall_s=wall_b,
                            packets_per_s=sum(npk) / (mc_ms_b * 1e-3),
                            voxel_updates_per_s=vox * args.steps / (heat_ms_b * 1e-3)),
               one_by_one=dict(mc_ms=mc_ms_1, heat_ms=heat_ms_1, wall_s=wall_1,
                               packets_per_s=sum(npk) / (mc_ms_1 * 1e-3),
                               voxel_updates_per_s=vox * args.steps / (heat_ms_1 * 1e-3)),
               same_fates=bool(same), max_abs_T_diff_first4=dT)
    json.dump(out, open(args.report, "w"), indent=1)
    log(json.dumps(out))


if __name__ == "__main__":
    main()
