"""Autotuner of the pair kernels' launch settings for one workload.

The settings change how the work is laid out on the GPU, never the neighbor
sets:
  threads per block and on-chip list length (lssph_set_tuning), particles
      per wavefront of the SPH_MODEL 6 kernels (lssph_set_wave_particles):
      the results are the same bit for bit;
  cell-list skin (lssph_set_skin) and cell columns per stencil reach
      (lssph_set_cell_split): cells of another width put the neighbors of a
      stencil row in another order, so the sums round differently (the FP64
      path's trajectories stay within 1e-12 of each other, the mixed path's
      within the FP32 rounding of its sums, i.e. within its own error
      against an FP64 run: up to ~5e-7 of a field's maximum over a few steps
      of a starting cavity flow).
`autotune` times a few graph-replayed steps per candidate on the caller's own
workload and returns the fastest.  No reference counterpart: the reference
has no launch settings.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field
from typing import Callable, Sequence

from .engine import LssphError, Simulation


@dataclass
class Setting:
    block: int        # threads per block of the pair kernels
    list_len: int     # on-chip neighbor-list entries (0: the size lssph_create chose)
    skin: float       # cell-list skin as a fraction of h_eff
    wave_particles: int = 0  # particles per wavefront, SPH_MODEL 6 (0: the library's default)
    cell_split: int = 0      # cell columns per stencil reach in x (0: the library's default)
    ms_per_step: float = float("nan")
    error: str = ""


@dataclass
class TuneResult:
    best: Setting
    default: Setting
    table: list = field(default_factory=list)

    def as_dict(self) -> dict:
        def d(s: Setting) -> dict:
            out = {"block": s.block, "list_len": s.list_len or "auto", "skin": s.skin,
                   "ms_per_step": round(s.ms_per_step, 4)}
            if s.wave_particles:
                out["wave_particles"] = s.wave_particles
            if s.cell_split:
                out["cell_split"] = s.cell_split
            if s.error:
                out["error"] = s.error
            return out
        return {"best": d(self.best), "default": d(self.default), "candidates": [d(s) for s in self.table]}


def _time_steps(sim: Simulation, steps: int) -> float:
    sim.step(1)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def autotune(make_sim: Callable[..., Simulation], steps: int = 5,
             blocks: Sequence[int] = (64, 128, 256), list_lens: Sequence[int] = (0,),
             skins: Sequence[float] = (0.04, 0.08, 0.16), wave_particles: Sequence[int] = (0,),
             splits: Sequence[int] = (0,), default: tuple = (128, 0, 0.08, 0, 0)) -> TuneResult:
    """make_sim(skin) -- make_sim(skin, cell_split) for a cell split other
    than 0, the library's default -- returns a fresh Simulation with those
    cells, uploaded and ready to step (both are fixed before the upload).
    Every combination of blocks x list_lens x skins x wave_particles x splits
    is timed over `steps` graph-replayed steps after one warm-up step; a
    combination the library rejects (an LDS limit, or a skin too narrow for
    the motion per step) is recorded with its error and skipped.  `default`
    is (block, list_len, skin, wave_particles, cell_split)."""
    table = []
    # list length 0 (the size lssph_create chose) must come first on each
    # fresh handle: lssph_set_tuning's 0 keeps the current length
    lens = sorted(set(list_lens), key=lambda n: (n != 0, n))
    for skin, split in itertools.product(skins, splits):
        sim = make_sim(skin, split) if split else make_sim(skin)
        try:
            for list_len, block, ppw in itertools.product(lens, blocks, wave_particles):
                s = Setting(block, list_len, skin, ppw, split)
                table.append(s)
                try:
                    sim.set_tuning(block=block, list_len=list_len)
                    if ppw:
                        sim.set_wave_particles(ppw)
                except LssphError as e:  # e.g. list_len x block beyond the LDS of a block
                    s.error = str(e)
                    continue
                try:
                    s.ms_per_step = _time_steps(sim, steps)
                except LssphError as e:  # e.g. the skin guard: this handle is done
                    s.error = str(e)
                    break
        finally:
            sim.close()
    ok = [s for s in table if not s.error]
    if not ok:
        raise RuntimeError("autotune: no candidate setting ran")
    best = min(ok, key=lambda s: s.ms_per_step)
    dflt = next((s for s in ok if (s.block, s.list_len, s.skin, s.wave_particles, s.cell_split)
                 == tuple(default) + (0,) * (5 - len(default))), best)
    return TuneResult(best, dflt, table)
