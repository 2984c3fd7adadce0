#!/usr/bin/env python3
"""Coupled photon-transport / heat-diffusion cycle throughput on one MI355X.

Headline leg (the configuration the metric is quoted on): three-layer skin
on a 200^3 grid over 2x2x1 cm, 1e8 photon packets from a Gaussian beam
(1/e^2 radius 0.15 cm) per cycle, 10,000 FP32 stencil steps per cycle with
the FP64 Arrhenius integral inside every step, ablation above 100 C and the
optics/property rebuild, repeated over coupling cycles.  A "step" of this
benchmark is one whole coupling cycle; --steps K times K cycles after
--warmup W cycles (the tissue is reset to its initial state in between).

Second leg (the metric's "coupled cycle time, denoised vs plain"): the same
skin on 256^3, 20 cycles of 5000 steps, plain (1e8 packets per cycle) against
the chosen form (1e6 packets + the denoiser, residual check with fallback);
an untimed third pass scores the denoised fluence against 1e8 packets.
--no-c4 skips it.

Prints one JSON line.  `value` is whole-job photon packets per second of the
headline leg (stencil work included); the FD voxel-updates per second, the
cycle time and the per-kernel rates (HIP-event times on the library's stream)
are reported beside it, with the roofline of the dominant kernel, the CPU
baseline (the FP64 oracle timed on this host on a bounded sample of the same
workload, threaded and single-core) and the "c4" object of the second leg.

Auxiliary legs for profiling (each prints its own JSON line):
  --leg stencil-state save PATH   run 3 headline cycles and save the tissue state
  --leg stencil-state run PATH    reload it and run --heat-steps steps per shape
                                  (--shapes tile:kz,...), after a 256 MiB copy
  --leg transport-state run PATH  transport runs on such a state (crater phase)
                                  at a fixed launch (--mc-launch, --mc-runs)
  --leg denoiser                  time forward passes and training steps of the
                                  denoiser at --grid (a multiple of 32)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tissue-ablation-mc_amd"))

HBM_PEAK_GBS = 8000.0                 # MI355X HBM3E
MFMA_BF16_PEAK_TFLOPS = 2500.0        # MI355X dense BF16 matrix peak
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


def cpu_baseline(prob, budget_s, threads=None):
    """FP64 oracle on this host: MC packets/s and stencil voxel-updates/s on a
    bounded sample, combined into the same cycle's packets/s."""
    import ctypes as C

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    from tam_amd import lib
    p = prob.params
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.lib()
    # the oracle's stencil runs OpenMP regions at the calling thread's count
    C.CDLL("libgomp.so.1").omp_set_num_threads(threads)
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


def c1_line(args):
    """The headline leg: coupling cycles of the 200^3 skin workload."""
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
    # the stencil's figure is the kernel of the launch shape that ran the most
    # timed steps (the stencil shape is re-tuned every cycle), at that shape's time
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
    # transport: one FP64 tally add per voxel visit outside the block's LDS
    # tally box goes to device memory; its ceiling (the device's global FP64
    # atomic rate) is measured on this device, and the VALU figure is the
    # algorithmic lane-ops per crossing / interaction over the lane-op peak
    atom_peak = run.grid.atomic_rate()
    atom = (T["tally_adds"] - T["box_adds"]) / mc_s
    mc_ops = (MC_OPS_CROSS * T["crossings"] + MC_OPS_INTER * T["interactions"]) / mc_s
    mc_roof = {"bound": "valu", "achieved": mc_ops / 1e12, "peak": LANE_OPS_PEAK / 1e12, "unit": "Tlane-op/s",
               "frac": mc_ops / LANE_OPS_PEAK,
               "ops_basis": f"{MC_OPS_CROSS:g} lane-ops per crossing + {MC_OPS_INTER:g} per interaction",
               "atomic": {"achieved": atom / 1e9, "peak_measured": atom_peak / 1e9, "unit": "Gatomic/s",
                          "frac": atom / atom_peak},
               "tally_adds_per_packet": T["tally_adds"] / T["packets"],
               "device_adds_per_packet": (T["tally_adds"] - T["box_adds"]) / T["packets"],
               "lds_box_adds_per_packet": T["box_adds"] / T["packets"]}
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
                             f"the transport kernel ({mc_s / wall:.0%}), VALU basis with the device-atomic "
                             "figure beside it")
    if not args.no_cpu:
        cb = cpu_baseline(prob, args.cpu_seconds)
        one = cpu_baseline(prob, max(2.0, args.cpu_seconds / 3), threads=1)
        cb["single_core"] = {k: one[k] for k in ("value", "packets_per_s", "voxel_updates_per_s", "cycle_s",
                                                 "sample")}
        line["cpu_baseline"] = cb
    run.close()
    return line


def c4_leg(args, log):
    """Denoised vs plain coupled cycles on the 256^3 skin (20 cycles of 5000
    steps by default), with the accuracy of the chosen form."""
    import numpy as np

    from tam_amd import configs, coupled
    from tam_amd.denoise import Denoiser

    n = args.c4_grid
    prob = configs.config4_tissue(n=(n, n, n), packets=int(1e8), steps=args.c4_steps, cycles=args.c4_cycles)
    dn = Denoiser((n, n, n), capacity=0)
    dn.set_params(np.fromfile(args.dn_weights, dtype=np.float32))
    kw = dict(cycles=args.c4_cycles, steps=args.c4_steps, hi=int(1e8), lo=int(1e6), tol=0.02)
    plain = coupled.run_form("plain", prob, **kw)
    log(f"c4 plain: {plain['wall_s']:.2f} s")
    chosen = coupled.run_form("chosen", prob, dn=dn, **kw)
    log(f"c4 chosen: {chosen['wall_s']:.2f} s")
    ev = coupled.run_form("chosen-eval", prob, dn=dn, evaluate=True, **kw)
    dn.close()
    rep = coupled.compare(prob, plain, chosen, ev)
    # the noise floor of the accuracy figures: a second plain run on other
    # Philox streams, compared with the first the same way (untimed)
    floor = None
    if not args.c4_no_floor:
        plain2 = coupled.run_form("plain", prob, seed=12, tune=False, **kw)
        floor = coupled.compare(prob, plain, plain2)
        log(f"c4 plain, second seed: peak T difference {floor['peak_T_error']:.3f} K")
    # per-cycle time split: transport (+ denoiser) vs the stencil both forms run
    heat_p = sum(c["heat_ms"] for c in plain["cycles"]) / 1e3
    heat_c = sum(c["heat_ms"] for c in chosen["cycles"]) / 1e3
    m_p, m_c = plain["wall_s"] - heat_p, chosen["wall_s"] - heat_c
    out = {"workload": f"three-layer skin {n}^3 over 2x2x1 cm, {args.c4_cycles} cycles x {args.c4_steps} FP32 "
                       "stencil steps (FP64 Arrhenius each step), ablation/rebuild; plain 1e8 packets/cycle vs "
                       "chosen 1e6 packets + denoiser, residual check vs an independent 1e6 run, fallback to "
                       "1e8 above 2%",
           "plain_s": rep["plain_s"], "chosen_s": rep["chosen_s"], "speedup": rep["speedup"],
           "fallbacks": rep["fallbacks"],
           "median_fluence_error": rep["median_fluence_error_denoised"],
           "median_fluence_error_plain_1e6": rep["median_fluence_error_plain_lo"],
           "peak_T_error_K": rep["peak_T_error"],
           "peak_T_basis": "max over cycles of |peak tissue T before ablation, plain - chosen|",
           "p99_abs_T_diff_tissue_K": rep["p99_abs_T_diff_tissue"],
           "max_abs_T_diff_tissue_K": rep["max_abs_T_diff_tissue"],
           "crater_cm3": [rep["crater_cm3_plain"], rep["crater_cm3_chosen"]],
           "crater_rel_diff": rep["crater_rel_diff"],
           "accuracy_met": {"fluence_le_2pct": rep["median_fluence_error_denoised"] <= 0.02,
                            "peak_T_le_0.5K": rep["peak_T_error"] <= 0.5},
           "split_s": {"plain_transport": m_p, "plain_stencil": heat_p, "chosen_transport_denoise": m_c,
                       "chosen_stencil": heat_c}}
    # 20x needs (m_p + S) / (m_c + S) >= 20 with S the stencil time both run:
    # S <= (m_p - 20 m_c) / 19, negative when m_p / m_c < 20 (no stencil speed reaches it)
    out["stencil_bound_for_20x_s"] = (m_p - 20.0 * m_c) / 19.0
    out["ratio_with_zero_stencil_time"] = m_p / m_c
    split = [c.get("split_ms", {}) for c in chosen["cycles"]]
    out["chosen_ms_per_cycle"] = {k: float(np.mean([q.get(k, 0.0) for q in split]))
                                  for k in ("independent_run_and_blocks", "transport", "denoise_and_blocks",
                                            "residual", "forward_ms")}
    out["chosen_ms_per_cycle"]["stencil"] = 1e3 * heat_c / len(chosen["cycles"])
    out["plain_ms_per_cycle"] = {"transport": 1e3 * m_p / len(plain["cycles"]),
                                 "stencil": 1e3 * heat_p / len(plain["cycles"])}
    if floor is not None:
        out["noise_floor_plain_vs_plain"] = {
            "peak_T_error_K": floor["peak_T_error"], "p99_abs_T_diff_tissue_K": floor["p99_abs_T_diff_tissue"],
            "max_abs_T_diff_tissue_K": floor["max_abs_T_diff_tissue"], "crater_rel_diff": floor["crater_rel_diff"],
            "basis": "two plain 1e8-packet runs on different Philox streams"}
    return out


def stencil_state_leg(args, log):
    """Save a coupled-cycle tissue state, or replay heat steps from one, so a
    counter run of the heat step at a late-cycle state does not have to run
    the cycles that build it."""
    import numpy as np

    from tam_amd import configs, lib
    from tam_amd.cycle import CycleRunner

    fields = dict(temp=lib.FIELD_TEMP, omega=lib.FIELD_TISSUE, thres=lib.FIELD_THRESTIME, kappa=lib.FIELD_KAPPA,
                  density=lib.FIELD_DENSITY, heatcap=lib.FIELD_HEATCAP, rhokap=lib.FIELD_RHOKAP,
                  jmean=lib.FIELD_JMEAN, water=lib.FIELD_WATER, q=lib.FIELD_Q)
    n = args.grid
    prob = configs.config1(n=(n, n, n), packets=int(args.packets), steps=10000, cycles=3)
    mode, path = args.state
    if mode == "save":
        run = CycleRunner(prob, seed=2024)
        for c in range(3):
            run.run_cycle()
            log("cycle", c, "done")
        g = run.grid
        g.normalise(prob.scale)
        np.savez(path, **{k: g.download(v) for k, v in fields.items()})
        g.close()
        return {"leg": "stencil-state", "saved": path, "grid": n, "cycles": 3}
    d = np.load(path)
    shapes = [tuple(int(x) for x in q.split(":")) for q in args.shapes.split(",")]
    res = {"leg": "stencil-state", "grid": n, "steps": args.heat_steps, "us_per_step": {}}
    for k_shape, (tile, kz) in enumerate(shapes):
        g = lib.TissueGrid(prob.params)
        g.set_optics(prob.mat, prob.table, d["rhokap"])          # removed voxels: rhokap 0
        g.set_thermal(d["temp"], d["kappa"], d["density"], d["heatcap"])
        for k in ("omega", "thres", "water", "q", "jmean"):
            g.upload(fields[k], d[k])
        if k_shape == 0:
            res["copy_gbps"] = g.copy_bandwidth(256 << 20, reps=3)
        g.set_heat_launch(tile, kz)
        done, ms = 0, 0.0
        while done < args.heat_steps:
            k = min(50, args.heat_steps - done)
            g.heat_run(k)
            ms += g.timing().heat_ms
            done += k
            log(f"tile {tile} kz {kz}, {done} steps")
        res["us_per_step"][f"{SHAPES[tile][0]}/kz{kz}"] = ms / done * 1e3
        g.close()
    return res


def transport_state_leg(args, log):
    """Transport on a saved coupled-cycle tissue state (the crater phase) at a
    fixed launch shape, --mc-runs runs of --packets packets, so a counter run
    of the transport kernel sees the state the timed cycles spend most of
    their transport time in.  The first run places the LDS tally box at the
    top face (no path yet); later runs place it from the run before."""
    import numpy as np

    from tam_amd import configs, lib

    n = args.grid
    prob = configs.config1(n=(n, n, n), packets=int(args.packets), steps=10000, cycles=3)
    d = np.load(args.state[1])
    g = lib.TissueGrid(prob.params)
    g.set_optics(prob.mat, prob.table, d["rhokap"])              # removed voxels: rhokap 0
    g.set_thermal(d["temp"], d["kappa"], d["density"], d["heatcap"])
    bpc, batch = (int(x) for x in args.mc_launch.split(":"))
    g.set_mc_launch(bpc, batch)
    g.set_mc_box(True)
    # a streaming copy of known size (256 MiB each way): the calibration of
    # the byte counters in a counter run of this leg
    copy_gbps = g.copy_bandwidth(256 << 20, reps=3)
    runs = []
    for r in range(args.mc_runs):
        st = g.mc_run(int(args.packets), seed=2024, stream=r)
        npk = float(args.packets)
        runs.append({"kernel_ms": st.kernel_ms, "packets_per_s": npk / (st.kernel_ms * 1e-3),
                     "device_adds_per_packet": (st.tally_adds - st.box_adds) / npk,
                     "box_adds_per_packet": st.box_adds / npk, "crossings_per_packet": st.crossings / npk,
                     "interactions_per_packet": st.interactions / npk, "box": g.mc_box()})
        log(f"transport run {r}: {runs[-1]['kernel_ms']:.1f} ms, {runs[-1]['device_adds_per_packet']:.1f} device "
            f"+ {runs[-1]['box_adds_per_packet']:.1f} box adds per packet")
    g.close()
    return {"leg": "transport-state", "grid": n, "packets": int(args.packets), "blocks_per_cu": bpc,
            "rng_batch": batch, "copy_gbps": copy_gbps, "runs": runs}


def denoiser_leg(args, log):
    """Forward passes and training steps of the denoiser on random feature
    maps (the committed weights), timed by the network's own device timer."""
    import numpy as np

    from tam_amd import learned
    from tam_amd.denoise import Denoiser

    n = args.grid
    slots = args.dn_slots if args.dn_train_steps > 0 else 0
    dn = Denoiser((n, n, n), capacity=slots, seed=1)
    dn.set_params(np.fromfile(args.dn_weights, dtype=np.float32))
    rng = np.random.default_rng(0)
    for q in range(slots):
        f = rng.random((4, n, n, n)).astype(np.float32)
        dn.store_host(q, f, 1.01 * f[0])
    f = rng.random((4, n, n, n)).astype(np.float32)
    flop = learned.conv_flops((n, n, n))
    fwd = []
    for _ in range(args.dn_forward):
        dn.forward(f)
        fwd.append(dn.last_ms())
        log(f"forward {n}^3: {fwd[-1]:.3f} ms")
    res = {"leg": "denoiser", "grid": n, "forward_ms": min(fwd) if fwd else None, "flop_forward": flop,
           "mfma_peak_tflops": MFMA_BF16_PEAK_TFLOPS}
    if fwd:
        res["forward_tflops"] = flop / (min(fwd) * 1e-3) / 1e12
        res["forward_frac"] = res["forward_tflops"] / MFMA_BF16_PEAK_TFLOPS
    if args.dn_train_steps > 0:
        dn.train(slots, args.dn_train_steps, lr=1e-4, seed=0)
        ms = dn.last_ms() / args.dn_train_steps
        res.update(train_ms_per_step=ms, train_tflops=3 * flop / (ms * 1e-3) / 1e12)
        res["train_frac"] = res["train_tflops"] / MFMA_BF16_PEAK_TFLOPS
        log(f"training {n}^3: {ms:.3f} ms/step")
    dn.close()
    return res


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
    ap.add_argument("--no-c4", action="store_true", help="skip the denoised-vs-plain leg")
    ap.add_argument("--c4-grid", type=int, default=256)
    ap.add_argument("--c4-cycles", type=int, default=20)
    ap.add_argument("--c4-steps", type=int, default=5000)
    ap.add_argument("--c4-no-floor", action="store_true", help="skip the second plain run (accuracy noise floor)")
    ap.add_argument("--leg", default="main", choices=("main", "stencil-state", "transport-state", "denoiser"))
    ap.add_argument("--state", nargs=2, metavar=("save|run", "PATH"), default=None)
    ap.add_argument("--shapes", default="4:8", help="stencil-state leg: comma list of tile:kz")
    ap.add_argument("--mc-launch", default="4:32", help="transport-state leg: blocks per CU:random-phase batch")
    ap.add_argument("--mc-runs", type=int, default=3, help="transport-state leg: transport runs")
    ap.add_argument("--dn-weights", default=os.path.join(ROOT, "tissue-ablation-mc_amd", "weights",
                                                         "denoiser_c3.f32"), help="denoiser parameter vector")
    ap.add_argument("--dn-forward", type=int, default=5)
    ap.add_argument("--dn-train-steps", type=int, default=200)
    ap.add_argument("--dn-slots", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.leg == "stencil-state":
        if args.state is None or args.state[0] not in ("save", "run"):
            ap.error("--leg stencil-state needs --state save|run PATH")
        print(json.dumps(stencil_state_leg(args, log)), flush=True)
        return
    if args.leg == "transport-state":
        if args.state is None or args.state[0] != "run":
            ap.error("--leg transport-state needs --state run PATH (a state saved by the stencil-state leg)")
        print(json.dumps(transport_state_leg(args, log)), flush=True)
        return
    if args.leg == "denoiser":
        print(json.dumps(denoiser_leg(args, log)), flush=True)
        return
    line = c1_line(args)
    log(f"c1 leg: {line['value']:.4g} photons/s, {line['cycle_time_s']:.3f} s per cycle")
    if not args.no_c4:
        line["c4"] = c4_leg(args, log)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
