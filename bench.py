#!/usr/bin/env python3
"""Coupled photon-transport / heat-diffusion cycle throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): three-layer
skin on a 200^3 grid over 2x2x1 cm, 1e8 photon packets from a Gaussian beam
(1/e^2 radius 0.15 cm) per cycle, 10,000 FP32 stencil steps per cycle with
the FP64 Arrhenius integral inside every step, ablation above 100 C and the
optics/property rebuild, repeated over coupling cycles.  A "step" of this
benchmark is one whole coupling cycle; --steps K times K cycles after
--warmup W cycles (the tissue is reset to its initial state in between).

Prints one JSON line.  `value` is whole-job photon packets per second over
the timed cycles (stencil work included); the FD voxel-updates per second,
the cycle time and the per-kernel rates (HIP-event times on the library's
stream) are reported beside it, with the roofline of the dominant kernel and
the CPU baseline (the FP64 oracle, OpenMP, timed on this host on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tissue-ablation-mc_amd"))

HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E
LANE_OPS_PEAK = 256 * 64 * 2.4e9      # 256 CU x 64 lanes x 2.4 GHz (non-packed FP32 lane-ops/s)
# algorithmic bytes per voxel-update: T read + write 8, source 4, material
# class 1, and with damage integrated in every step the FP64 Omega read +
# write 16 (the reference's Arrhenius pass touches Omega for every voxel)
STENCIL_BYTES_BASE = 13.0
STENCIL_BYTES_DAMAGE = 16.0
MC_OPS_CROSS, MC_OPS_INTER = 16.0, 64.0   # algorithmic FP32 lane-ops per crossing / interaction
# heat launch shapes (tam_set_heat_launch tile ids): name, kernel, steps per launch
SHAPES = [("lds-64x4", "stencil_kernel<64,4>", 1), ("lds-32x8", "stencil_kernel<32,8>", 1),
          ("lds-64x8", "stencil_kernel<64,8>", 1), ("direct-64x4", "stencil_direct_kernel<64,4>", 1),
          ("direct-32x8", "stencil_direct_kernel<32,8>", 1), ("direct-64x8", "stencil_direct_kernel<64,8>", 1),
          ("two-step-32x8", "stencil_t2_kernel<32,8>", 2), ("two-step-32x16", "stencil_t2_kernel<32,16>", 2),
          ("two-step-64x8", "stencil_t2_kernel<64,8>", 2)]


def measured_traffic(n, shape):
    """HBM bytes per launch of the heat kernel of launch shape `shape` from the
    committed counter summary (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each
    scaled by the same counter's reading for the streaming copy of known size),
    or None where none was recorded for that shape."""
    path = os.path.join(ROOT, "profiles", f"stencil_traffic_{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get("by_shape", {}).get(shape)
    return rec.get("bytes_per_launch") if rec else None


def cpu_baseline(prob, budget_s):
    """FP64 oracle on this host: MC packets/s and stencil voxel-updates/s on a
    bounded sample, combined into the same cycle's packets/s."""
    import ctypes as C

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    from tam_amd import lib
    p = prob.params
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    mc = O.MCProblem(p.nx, p.ny, p.nz, p.xmax, p.ymax, p.zmax, prob.mat, prob.table,
                     n_ext=(0, 0, 0, 0, p.n_ext_zlo, p.n_ext_zhi), beam=p.beam, radius=p.beam_radius,
                     rng=0, seed=5)
    npk, t_mc, batch = 0, 0.0, 20000
    while t_mc < 0.6 * budget_s and npk < 50_000_000:
        t0 = time.perf_counter()
        mc.run(batch, p0=npk, nthreads=threads)
        t_mc += time.perf_counter() - t0
        npk += batch
        batch = min(batch * 2, 2_000_000)
    pk_rate = npk / t_mc
    # heat: the oracle's reference-order FP64 step on the full grid
    hs = (p.nz + 2, p.ny + 2, p.nx + 2)
    T = np.ascontiguousarray(prob.temp.copy())
    coeff = np.zeros(hs)
    coeff[1:-1, 1:-1, 1:-1] = p.delt / (prob.density * prob.heatcap)[1:-1, 1:-1, 1:-1]
    jm = np.ascontiguousarray(np.full((p.nz, p.ny, p.nx), 1e6))
    Q = np.zeros((p.nz, p.ny, p.nx))
    arrs = [np.ascontiguousarray(a) for a in (prob.kappa, prob.density, prob.heatcap, coeff)]
    h = O.Heat()
    h.nx, h.ny, h.nz = p.nx, p.ny, p.nz
    h.dx, h.dy, h.dz, h.delt, h.dscale = p.dx_m, p.dy_m, p.dz_m, p.delt, p.dscale
    h.temp = T.ctypes.data
    h.kappa, h.density, h.heatcap, h.coeff = [a.ctypes.data for a in arrs]
    h.Q, h.jmean = Q.ctypes.data, jm.ctypes.data
    h.volumeVoxel, h.massVoxel, h.QVapor = p.volume_voxel, p.mass_voxel, p.qvapor
    h.robin_top, h.h_top, h.T_amb = p.robin_top, p.h_top, p.t_amb
    steps, t_fd = 0, 0.0
    while t_fd < 0.4 * budget_s and steps < 1000:
        t0 = time.perf_counter()
        O.lib().or_heat_steps(C.byref(h), 1, None)
        t_fd += time.perf_counter() - t0
        steps += 1
    vu_rate = steps * p.nx * p.ny * p.nz / t_fd
    cyc = prob.packets / pk_rate + prob.steps * p.nx * p.ny * p.nz / vu_rate
    return {"value": prob.packets / cyc, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 oracle: {npk} packets of this tissue ({pk_rate:.4g} packets/s) and "
                       f"{steps} full-grid heat steps ({vu_rate:.4g} voxel-updates/s) on {threads} "
                       f"threads, combined into one cycle of {prob.packets:.3g} packets + "
                       f"{prob.steps} steps (extrapolated)"),
            "packets_per_s": pk_rate, "voxel_updates_per_s": vu_rate, "cycle_s": cyc}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="coupling cycles timed")
    ap.add_argument("--warmup", type=int, default=1, help="untimed cycles first")
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--heat-steps", type=int, default=10000)
    ap.add_argument("--grid", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from tam_amd import configs
    from tam_amd.cycle import CycleRunner

    n = args.grid
    prob = configs.config1(n=(n, n, n), packets=int(args.packets), steps=args.heat_steps,
                           cycles=args.steps)
    run = CycleRunner(prob, seed=2024)
    # measured streaming-copy bandwidth (256 MiB each way): the reference for
    # the stencil's roofline and the calibration of the byte counters
    copy_gbps = run.grid.copy_bandwidth(256 << 20, reps=20)
    if args.warmup > 0:
        run.run(args.warmup)
        run.reset()
    wall = run.run(args.steps)
    T = run.totals
    p = prob.params
    nvox = p.nx * p.ny * p.nz
    cycles = args.steps
    value = T["packets"] / wall
    vu = T["voxel_updates"] / wall
    mc_s, heat_s = T["mc_ms"] / 1e3, T["heat_ms"] / 1e3
    # the roofline is the kernel of the launch shape that ran the most timed
    # steps (the stencil shape is re-tuned every cycle), at that shape's time
    (tile, kz), (shape_ms, shape_steps) = max(run.heat_by_shape.items(), key=lambda kv: kv[1][1])
    step_ms = shape_ms / max(1, shape_steps)
    shape, kernel, per_launch = SHAPES[tile]
    launch_ms = step_ms * per_launch
    # algorithmic bytes per voxel-update x the voxel-updates of one launch (a
    # two-step launch updates every voxel twice); the traffic figure beside it
    # is what the launch moves (the two-step kernel keeps the intermediate
    # field on chip, so it moves less than the algorithmic bytes)
    per_vu = STENCIL_BYTES_BASE + (STENCIL_BYTES_DAMAGE if p.damage_every_step else 0.0)
    units = per_launch * nvox
    stencil = {"bound": "hbm", "achieved": per_vu * units / (launch_ms * 1e-3) / 1e9,
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_voxel_update": per_vu,
               "voxel_updates_per_launch": units, "steps_per_launch": per_launch, "launch_us": launch_ms * 1e3,
               "step_us": step_ms * 1e3,
               "bytes_basis": "T read+write 8, source 4, class 1" +
                              (", Omega read+write 16 (damage in every step)" if p.damage_every_step else "")}
    stencil["frac"] = stencil["achieved"] / stencil["peak"]
    stencil["measured_copy_gbps"] = copy_gbps
    stencil["traffic"] = measured_traffic(n, shape)
    # transport: bound by the global FP64 atomic rate of its tally (one add per
    # voxel visit outside the block's LDS tally box), whose ceiling is measured
    # on this device; deposits into the LDS box are counted apart
    atom_peak = run.grid.atomic_rate()
    atom = (T["tally_adds"] - T["box_adds"]) / mc_s
    mc_ops = (MC_OPS_CROSS * T["crossings"] + MC_OPS_INTER * T["interactions"]) / mc_s
    mc_roof = {"bound": "atomic", "achieved": atom / 1e9, "peak": atom_peak / 1e9, "unit": "Gatomic/s",
               "frac": atom / atom_peak, "tally_adds_per_packet": T["tally_adds"] / T["packets"],
               "lds_box_adds_per_packet": T["box_adds"] / T["packets"],
               "valu": {"achieved": mc_ops / 1e12, "peak": LANE_OPS_PEAK / 1e12, "unit": "Tlane-op/s",
                        "frac": mc_ops / LANE_OPS_PEAK}}
    roofline = dict(stencil, kernel=f"{kernel} ({shape}, kz {kz})",
                    steps_at_this_shape=f"{shape_steps} of {int(T['heat_steps'])}",
                    shapes_used={f"{SHAPES[t][0]}/kz{k}": v[1] for (t, k), v in run.heat_by_shape.items()})
    line = {
        "metric": "photons/s + FD voxel-steps/s; coupled cycle time, denoised vs plain @<=2% err",
        "value": value, "unit": "photons/s", "n_gpus": 1, "steps": cycles, "warmup": args.warmup,
        "ms_per_step": wall / cycles * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"three-layer skin {n}^3 over 2x2x1 cm, Gaussian w=0.15 cm, "
                               f"{prob.packets:.3g} packets + {prob.steps} FP32 stencil steps "
                               "(FP64 Arrhenius each step) + ablation/rebuild per cycle",
                   "global_batch": int(prob.packets), "seq_len": prob.steps, "parallelism": "none",
                   "grid": [n, n, n], "packets_per_cycle": int(prob.packets),
                   "heat_steps_per_cycle": prob.steps, "tally": "f64", "damage": "f64"},
        "voxel_updates_per_s": vu,
        "cycle_time_s": wall / cycles,
        "mc_packets_per_s_kernel": T["packets"] / mc_s,
        "stencil_voxel_updates_per_s_kernel": T["voxel_updates"] / heat_s,
        "time_split_s": {"mc": mc_s, "heat": heat_s, "rebuild": T["rebuild_ms"] / 1e3,
                         "normalise": T["normalise_ms"] / 1e3, "wall": wall},
        "mc_events": {"crossings_per_packet": T["crossings"] / T["packets"],
                      "interactions_per_packet": T["interactions"] / T["packets"]},
        "roofline": roofline,
        "roofline_mc": mc_roof,
        "launch": {"stencil_tile": int(run.grid.timing().tile), "stencil_kz": int(run.grid.timing().kz),
                   "class_mode": bool(run.grid.timing().class_mode),
                   "mc_blocks_per_cu": run.tuned_mc[0] if run.tuned_mc else None,
                   "mc_rng_batch": run.tuned_mc[1] if run.tuned_mc else None},
        "cpu_baseline": None,
    }
    line["roofline_note"] = (f"roofline: the HBM-bound stencil ({heat_s / wall:.0%} of the cycle); roofline_mc: "
                             f"the transport kernel ({mc_s / wall:.0%}), bound by the device's FP64 atomic rate")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(prob, args.cpu_seconds)
    run.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
