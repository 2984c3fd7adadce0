"""The coupled ablation run in two forms (mcpolar.f90:148-186 with the
transport replaced or not):

  plain   `hi` packets per cycle;
  chosen  `lo` packets + the denoiser.  A residual check compares 16^3-block
          sums of the denoised fluence with an independent `lo`-packet run
          (block sums on the device) and falls back to a `hi`-packet plain
          run for the cycle when the median block residual exceeds `tol`.

Both forms then normalise, run the same FP32 stencil steps (FP64 Arrhenius
inside every step) and the ablation / property rebuild.  An evaluation pass
repeats the chosen form and also runs `hi` packets on its tissue state in
every cycle (not timed) to score the denoised fluence.

compare() turns two runs into the accuracy report: the median per-voxel
fluence error, the per-cycle peak temperature before ablation of each form
and its largest difference, the 99th percentile and maximum |dT| over tissue
that survives in both forms, and the crater volumes.
"""
import time

import numpy as np

from . import learned, lib

BLOCK = 16


def block_residual(A, B, frac=1e-2):
    """Median over blocks (B's sum above frac of the largest) of
    |A - B| / B, for block sums A (denoised) and B (independent run)."""
    sel = B > frac * B.max()
    return float(np.median(np.abs(A[sel] - B[sel]) / B[sel]))


def run_form(form, prob, cycles, steps, hi, lo, tol=0.02, dn=None, evaluate=False, tune=True, seed=11,
             retune=(1, 3, 7, 15), log=None):
    """One form over `cycles` coupling cycles on a fresh grid of `prob`.
    Returns a record: per-cycle details, fallbacks, the wall time of the
    timed work (evaluation runs and downloads excluded), final T and mat."""
    p = prob.params
    g = prob.grid()
    if tune:
        g.tune_heat(reps=3)
        g.tune_mc(min(2_000_000, hi))
    vcm = p.volume_voxel * 1e6
    rec = dict(form=form, cycles=[], fallbacks=0)
    g.synchronize()
    t_start = time.perf_counter()
    t_eval = 0.0                      # evaluation-only work and logging, not timed
    for c in range(cycles):
        cyc = dict(cycle=c)
        split = {}
        t_ph = time.perf_counter()
        if form == "plain":
            st = g.mc_run(hi, seed=seed, stream=c)
            npk = hi
            split["transport"] = 1e3 * (time.perf_counter() - t_ph)
        else:
            if evaluate:
                t0 = time.perf_counter()
                g.mc_run(hi, seed=seed, stream=5000 + c)
                ref = g.download(lib.FIELD_TALLY) / (hi * vcm)
                t_eval += time.perf_counter() - t0
            t_ph = time.perf_counter()
            g.mc_run(lo, seed=seed, stream=3000 + c)
            B = g.tally_blocks(BLOCK)
            split["independent_run_and_blocks"] = 1e3 * (time.perf_counter() - t_ph)
            if evaluate:
                t0 = time.perf_counter()
                ind = g.download(lib.FIELD_TALLY)
                t_eval += time.perf_counter() - t0
            t_ph = time.perf_counter()
            st = g.mc_run(lo, seed=seed, stream=c)
            t1 = time.perf_counter()
            dn.apply(g, lo)
            A = g.tally_blocks(BLOCK)
            t2 = time.perf_counter()
            resid = block_residual(A, B)
            split.update(transport=1e3 * (t1 - t_ph), denoise_and_blocks=1e3 * (t2 - t1),
                         residual=1e3 * (time.perf_counter() - t2), forward_ms=dn.last_ms())
            cyc["residual"] = resid
            npk = lo
            if evaluate:
                t0 = time.perf_counter()
                den = g.download(lib.FIELD_TALLY)
                phi_d, phi_i = den / (lo * vcm), ind / (lo * vcm)
                cyc["fluence_error_denoised"] = learned.fluence_error(phi_d, ref)
                cyc["fluence_error_plain_lo"] = learned.fluence_error(phi_i, ref)
                cyc["fluence_bias_denoised"] = learned.fluence_bias(phi_d, ref)
                t_eval += time.perf_counter() - t0
            if resid > tol:
                rec["fallbacks"] += 1
                t_ph = time.perf_counter()
                st = g.mc_run(hi, seed=seed, stream=c)
                split["fallback"] = 1e3 * (time.perf_counter() - t_ph)
                npk = hi
                cyc["fallback"] = True
        g.normalise(prob.power / (npk * p.volume_voxel))
        if tune and c in retune:
            # the damage window fills over the first cycles and the fastest
            # stencil shape moves with it: re-time the shapes on the state
            g.tune_heat(reps=3)
        g.heat_run(steps)
        g.synchronize()
        t0 = time.perf_counter()
        T = g.download(lib.FIELD_TEMP)[1:-1, 1:-1, 1:-1]
        m = g.download(lib.FIELD_MAT)
        tissue = m != len(prob.table)
        cyc["peak_T_before_ablation"] = float(T[tissue].max())
        t_eval += time.perf_counter() - t0
        g.rebuild()
        g.synchronize()
        t0 = time.perf_counter()
        m = g.download(lib.FIELD_MAT)
        tm = g.timing()
        cyc.update(npk=npk, removed=int((m == len(prob.table)).sum()), heat_ms=tm.heat_ms, split_ms=split,
                   heat_shape=[int(tm.tile), int(tm.kz)], mc_ms=st.kernel_ms)
        rec["cycles"].append(cyc)
        if log:
            log(form, cyc)
        t_eval += time.perf_counter() - t0
    g.synchronize()
    rec["wall_s"] = time.perf_counter() - t_start - t_eval
    rec["T"] = g.download(lib.FIELD_TEMP)[1:-1, 1:-1, 1:-1]
    rec["mat"] = g.download(lib.FIELD_MAT)
    g.close()
    return rec


def compare(prob, plain, chosen, ev=None):
    """Accuracy of the chosen form against the plain one."""
    nx, ny, nz = prob.params.nx, prob.params.ny, prob.params.nz
    vox_cm3 = (2 * prob.params.xmax / nx) * (2 * prob.params.ymax / ny) * (2 * prob.params.zmax / nz)
    air = len(prob.table)
    tp, tc = plain["mat"] != air, chosen["mat"] != air        # voxels still tissue
    both = tp & tc
    dT = np.abs(plain["T"] - chosen["T"])[both]
    pk_p = [c["peak_T_before_ablation"] for c in plain["cycles"]]
    pk_c = [c["peak_T_before_ablation"] for c in chosen["cycles"]]
    out = dict(plain_s=plain["wall_s"], chosen_s=chosen["wall_s"], speedup=plain["wall_s"] / chosen["wall_s"],
               fallbacks=chosen["fallbacks"],
               peak_T_before_ablation_plain=pk_p, peak_T_before_ablation_chosen=pk_c,
               peak_T_error=float(np.max(np.abs(np.subtract(pk_p, pk_c)))),
               p99_abs_T_diff_tissue=float(np.percentile(dT, 99)), max_abs_T_diff_tissue=float(dT.max()),
               median_abs_T_diff_tissue=float(np.median(dT)),
               crater_cm3_plain=float((~tp).sum() * vox_cm3), crater_cm3_chosen=float((~tc).sum() * vox_cm3),
               crater_voxels_symdiff=int(((~tp) != (~tc)).sum()))
    out["crater_rel_diff"] = (out["crater_cm3_chosen"] - out["crater_cm3_plain"]) / max(out["crater_cm3_plain"],
                                                                                        1e-300)
    if ev is not None:
        e = [c["fluence_error_denoised"][0] for c in ev["cycles"]]
        e1 = [c["fluence_error_plain_lo"][0] for c in ev["cycles"]]
        b = [c["fluence_bias_denoised"] for c in ev["cycles"]]
        out.update(median_fluence_bias_denoised=float(np.median([x[0] for x in b])),
                   sum_fluence_bias_denoised=float(np.median([x[1] for x in b])))
        out.update(median_fluence_error_denoised=float(np.median(e)),
                   max_cycle_fluence_error_denoised=float(np.max(e)),
                   median_fluence_error_plain_lo=float(np.median(e1)))
    return out
