"""One MC-heat coupling cycle, as the body of mcpolar's time loop
(mcpolar.f90:148-186): photon transport, normalisation, `loops` heat steps,
Arrhenius damage (at the reference cadence unless it is integrated inside
every step), and the ablation / property rebuild."""
import time as _time

from . import lib


class CycleRunner:
    """graph=True runs each cycle as one replay of the captured HIP graph
    (tam_cycle); graph=False issues the same steps one C-ABI call at a time."""

    def __init__(self, prob, seed=0, tune=True, graph=False):
        self.prob = prob
        self.seed = seed
        self.graph = graph
        self.grid = prob.grid()
        # launch shapes picked by timing the candidates on this grid
        self.tuned = self.grid.tune_heat(reps=5) if tune else None
        self.tuned_mc = self.grid.tune_mc(min(int(prob.packets), 5_000_000)) if tune else None
        self.time = 0.0
        self.cycle = 0
        self.totals = dict(packets=0.0, crossings=0.0, interactions=0.0, tally_adds=0.0, box_adds=0.0,
                           mc_ms=0.0, heat_ms=0.0,
                           heat_steps=0, rebuild_ms=0.0, normalise_ms=0.0, arrhenius_ms=0.0,
                           voxel_updates=0.0)
        self.heat_by_shape = {}      # (tile, kz) -> [heat ms, steps] of the stepwise cycles

    def reset(self):
        """Restore the problem's initial tissue (optics, temperature, damage)."""
        p = self.prob
        self.grid.set_optics(p.mat, p.table, p.rhokap)
        self.grid.set_thermal(p.temp, p.kappa, p.density, p.heatcap)
        self.grid.upload(lib.FIELD_TISSUE, 0.0 * p.mat.astype(float))
        self.grid.upload(lib.FIELD_THRESTIME, 0.0 * (p.mat.astype(float)[None].repeat(3, 0)))
        self.time = 0.0
        self.cycle = 0
        for k in self.totals:
            self.totals[k] = 0.0
        self.heat_by_shape = {}

    def _retune(self):
        """The damage window and the source change from cycle to cycle, fastest
        early on: re-time the stencil shapes on the field of cycles 0, 1, 2, 4,
        8, 16, ... (long cycles only)."""
        c = self.cycle
        return self.tuned is not None and self.prob.steps >= 200 and (c & (c - 1)) == 0

    def run_cycle(self):
        if self.graph:
            return self._run_cycle_graph()
        p, g = self.prob, self.grid
        st = g.mc_run(p.packets, seed=self.seed, stream=self.cycle)
        g.normalise(p.scale)
        if self._retune():
            self.tuned = g.tune_heat(reps=3)
        g.heat_run(p.steps)
        self.time += p.steps * p.params.delt
        if not p.params.damage_every_step:
            g.arrhenius(p.params.delt, self.time)
        g.rebuild()
        t = g.timing()
        T = self.totals
        T["packets"] += st.packets
        T["crossings"] += st.crossings
        T["interactions"] += st.interactions
        T["tally_adds"] += st.tally_adds
        T["box_adds"] += st.box_adds
        T["mc_ms"] += st.kernel_ms
        T["heat_ms"] += t.heat_ms
        T["heat_steps"] += t.heat_steps
        acc = self.heat_by_shape.setdefault((int(t.tile), int(t.kz)), [0.0, 0])
        acc[0] += t.heat_ms
        acc[1] += t.heat_steps
        T["rebuild_ms"] += t.rebuild_ms
        T["normalise_ms"] += t.normalise_ms
        T["arrhenius_ms"] += 0.0 if p.params.damage_every_step else t.arrhenius_ms
        T["voxel_updates"] += float(p.steps) * p.params.nx * p.params.ny * p.params.nz
        self.cycle += 1
        return st

    def _run_cycle_graph(self):
        p, g = self.prob, self.grid
        if self._retune():
            self.tuned = g.tune_heat(reps=3)         # keeps the graph unless the shape changes
        arr = 0.0 if p.params.damage_every_step else p.params.delt
        st = g.cycle(p.packets, self.seed, self.cycle, p.scale, p.steps, arr_delt=arr)
        self.time += p.steps * p.params.delt
        t = g.timing()
        T = self.totals
        T["packets"] += st.packets
        T["crossings"] += st.crossings
        T["interactions"] += st.interactions
        T["tally_adds"] += st.tally_adds
        T["box_adds"] += st.box_adds
        T["cycle_ms"] = T.get("cycle_ms", 0.0) + t.cycle_ms
        T["voxel_updates"] += float(p.steps) * p.params.nx * p.params.ny * p.params.nz
        self.cycle += 1
        return st

    def run(self, cycles):
        t0 = _time.perf_counter()
        for _ in range(cycles):
            self.run_cycle()
        self.grid.synchronize()
        return _time.perf_counter() - t0

    def close(self):
        self.grid.close()
