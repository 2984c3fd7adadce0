"""Save a coupled-cycle tissue state, or replay heat steps from a saved one.

Counter runs (rocprofv3 --pmc) serialise every dispatch, so a profile of the
heat step at a late-cycle state cannot afford to run the cycles that build
that state.  This script splits the two:

    # 3 full c1 cycles at 200^3 (1e8 packets + 10 000 steps each), then save
    python tissue-ablation-mc_amd/stencil_state.py save /tmp/c1_state.npz --cycles 3

    # reload the state and run 200 heat steps (plus a 256 MiB streaming copy,
    # the byte-counter calibration) with a given launch shape
    python tissue-ablation-mc_amd/stencil_state.py run /tmp/c1_state.npz --tile 4 --kz 8

    # time several launch shapes (tile:kz), each from the saved state
    python tissue-ablation-mc_amd/stencil_state.py run /tmp/c1_state.npz --shapes 4:8,8:8,8:16

The saved fields are the temperature (with halo), Omega, the threshold times,
water, Q, jmean (the next cycle's source) and the rebuilt kappa, density,
heat capacity and rhokap (0 in removed voxels).
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from tam_amd import configs, lib  # noqa: E402
from tam_amd.cycle import CycleRunner  # noqa: E402

FIELDS = dict(temp=lib.FIELD_TEMP, omega=lib.FIELD_TISSUE, thres=lib.FIELD_THRESTIME, kappa=lib.FIELD_KAPPA,
              density=lib.FIELD_DENSITY, heatcap=lib.FIELD_HEATCAP, rhokap=lib.FIELD_RHOKAP, jmean=lib.FIELD_JMEAN,
              water=lib.FIELD_WATER, q=lib.FIELD_Q)


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=("save", "run"))
    ap.add_argument("path")
    ap.add_argument("--grid", type=int, default=200)
    ap.add_argument("--cycles", type=int, default=3)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--tile", type=int, default=4)
    ap.add_argument("--kz", type=int, default=8)
    ap.add_argument("--shapes", default=None, help="comma list of tile:kz (overrides --tile/--kz)")
    args = ap.parse_args()
    n = args.grid
    prob = configs.config1(n=(n, n, n), packets=int(1e8), steps=10000, cycles=args.cycles)
    if args.mode == "save":
        run = CycleRunner(prob, seed=2024)
        for c in range(args.cycles):
            run.run_cycle()
            log("cycle", c, "done")
        g = run.grid
        g.normalise(prob.scale)
        np.savez(args.path, **{k: g.download(v) for k, v in FIELDS.items()})
        log("saved", args.path)
        g.close()
        return
    d = np.load(args.path)
    shapes = ([tuple(int(x) for x in s.split(":")) for s in args.shapes.split(",")] if args.shapes
              else [(args.tile, args.kz)])
    for n_shape, (tile, kz) in enumerate(shapes):
        g = lib.TissueGrid(prob.params)
        g.set_optics(prob.mat, prob.table, d["rhokap"])          # removed voxels: rhokap 0
        g.set_thermal(d["temp"], d["kappa"], d["density"], d["heatcap"])
        for k in ("omega", "thres", "water", "q", "jmean"):
            g.upload(FIELDS[k], d[k])
        log("loaded", args.path)
        if n_shape == 0:
            log("copy GB/s", g.copy_bandwidth(256 << 20, reps=3))
        g.set_heat_launch(tile, kz)
        done = 0
        while done < args.steps:
            k = min(50, args.steps - done)
            g.heat_run(k)
            done += k
            log(f"tile {tile} kz {kz}, {k} steps: {g.timing().heat_ms / k * 1e3:.2f} us/step")
        g.close()


if __name__ == "__main__":
    main()
