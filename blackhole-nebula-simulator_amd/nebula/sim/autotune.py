"""Launch-shape autotuner for the mixed-precision hydro pass (no counterpart in the reference).

Two knobs change the speed of the density and force passes but not their physics:

- ``cell_factor``: the neighbour-grid cell edge in units of the mean search length. Smaller
  cells trim the candidate walk; larger cells mean fewer, longer cell ranges.
- ``tile``: the workgroup size (64, 128 or 256 threads, i.e. 4, 8 or 16 particle rows of 16
  lanes) of the lane-group kernels.

The tile changes nothing in the results (every row is summed by its own lane group). The cell
factor changes the grid, hence the order in which a row's pairs are visited, so mixed-precision
sums move within their rounding; the lists themselves and every FP64 quantity are unchanged.

``tune`` times one full hydro pass (grid, lists, h-rho solve, gradients, forces) of the live
state for every candidate and returns the fastest; ``apply`` makes it the state's setting.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, replace

import torch

from .. import _hip
from . import coupled
from .neighbours import permute_many

FACTORS = (0.75, 1.0, 1.25, 1.5)
TILES = (64, 128, 256)


@dataclass
class Choice:
    cell_factor: float
    tile: int
    ms: float
    table: dict  # (cell_factor, tile) -> milliseconds per hydro pass


def _time_pass(st, cfg, repeats: int) -> float:
    v = coupled.velocity(cfg, st.x, st.p)
    coupled._hydro(st, cfg, v, 0.0, st.a_hydro)  # warm-up: buffers sized, capacities settled
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        coupled._hydro(st, cfg, v, 0.0, st.a_hydro)
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def tune(st, cfg, factors=FACTORS, tiles=TILES, repeats: int = 3) -> Choice:
    """Fastest (cell_factor, tile) for the hydro pass of ``st`` (mixed precision only). The state
    is not changed; the library's tile is left at the winner."""
    if cfg.hydro.precision != "mixed":
        raise ValueError("the autotuner tunes the mixed-precision passes")
    table = {}
    for f in factors:
        trial = replace(st, cell_factor=float(f))
        order = coupled.grid_permutation(cfg, trial.x, trial.h, trial)
        x, p, u, m, h, a, acc = permute_many((st.x, st.p, st.u, st.mass, st.h, st.alpha, st.a_hydro), order)
        counts = None
        if st.row_counts is not None:
            n = int(st.x.shape[0])
            counts = permute_many((st.row_counts.view(torch.int64).view(n),), order)[0].view(torch.int32).view(n, 2)
        trial = replace(trial, x=x, p=p, u=u, mass=m, h=h, alpha=a, a_hydro=acc, row_counts=counts)
        for t in tiles:
            _hip.set_tile(t)
            table[(float(f), int(t))] = _time_pass(trial, cfg, repeats)
    (f, t), ms = min(table.items(), key=lambda kv: kv[1])
    _hip.set_tile(t)
    return Choice(f, t, ms, table)


def apply(st, choice: Choice):
    """The state with the chosen cell factor (used from the next grid on); sets the tile."""
    _hip.set_tile(choice.tile)
    return replace(st, cell_factor=float(choice.cell_factor))
