"""Neighbour lists, oracle side (test infrastructure only).

Restates the reference's cell-grid search ``find_neighbours_numba`` (neighbours.py:240-281),
``tighten`` (neighbours.py:306-323) and the O(N^2) definition ``brute_force_neighbours``
(neighbours.py:74-82). Lists are CSR, sorted, symmetric, including self.
"""

from dataclasses import dataclass

import numpy as np

from . import lib

R_KERN = 2.0


@dataclass
class Neighbours:
    """CSR lists: neighbours of i are index[offset[i]:offset[i+1]] (neighbours.py:22-36)."""

    offset: np.ndarray
    index: np.ndarray
    box: np.ndarray | None

    def of(self, i: int) -> np.ndarray:
        return self.index[self.offset[i] : self.offset[i + 1]]

    @property
    def counts(self) -> np.ndarray:
        return np.diff(self.offset)


def wrap(pos, box):
    """neighbours.py:46-50"""
    return pos if box is None else np.mod(pos, box)


def minimum_image(dx, box):
    """neighbours.py:39-43"""
    return dx if box is None else dx - box * np.round(dx / box)


def _box3(box):
    return None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()


def grid_parameters(wrapped: np.ndarray, h: np.ndarray, box_arr):
    """lo, inv_cell, ncell exactly as neighbours.py:248-265 chooses them."""
    if box_arr is None:
        lo = wrapped.min(axis=0)
        span = wrapped.max(axis=0) - lo
    else:
        lo = np.zeros(3)
        span = box_arr
    cell = max(0.5 * R_KERN * float(np.median(h)), float(np.max(span)) / 160.0, 1e-300)
    if box_arr is not None:
        ncell = np.maximum(np.floor(box_arr / cell).astype(np.int64), 1)
        inv_cell = ncell / box_arr
    else:
        ncell = np.maximum(np.floor(span / cell).astype(np.int64) + 1, 1)
        inv_cell = np.full(3, 1.0 / cell)
    return lib.f64(lo), lib.f64(inv_cell), lib.i64(ncell)


def find_neighbours_grid(pos, h, box=None) -> Neighbours:
    """neighbours.py:240-281: cell grid gather {j : r_ij < 2 h_i}, rows sorted, union with
    the transpose."""
    pos = lib.f64(pos)
    h = lib.f64(h)
    n = len(pos)
    box_arr = _box3(box)
    if box_arr is not None and np.any(R_KERN * h.max() >= 0.5 * box_arr):
        raise ValueError("kernel support exceeds half the periodic box: neighbours ambiguous")
    boxv = np.zeros(3) if box_arr is None else box_arr
    wrapped = lib.f64(wrap(pos, box_arr))
    lo, inv_cell, ncell = grid_parameters(wrapped, h, box_arr)
    total = int(np.prod(ncell))
    start = np.empty(total + 1, np.int64)
    order = np.empty(n, np.int64)
    lib.build_grid(lib.d(wrapped), n, lib.d(lo), lib.d(inv_cell), lib.i(ncell), lib.i(start),
                   lib.i(order))  # fmt: skip
    counts = np.empty(n, np.int64)
    dummy = np.empty(1, np.int64)
    args = (lib.d(wrapped), lib.d(h), n, lib.d(boxv), lib.d(lo), lib.d(inv_cell), lib.i(ncell),
            lib.i(start), lib.i(order), lib.i(counts))  # fmt: skip
    lib.gather(*args, lib.i(dummy), lib.i(dummy), 0)
    g_off = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=g_off[1:])
    g_idx = np.empty(max(int(g_off[-1]), 1), np.int64)
    lib.gather(*args, lib.i(g_off), lib.i(g_idx), 1)
    lib.sort_rows(lib.i(g_off), n, lib.i(g_idx))
    t_off = np.empty(n + 1, np.int64)
    t_idx = np.empty_like(g_idx)
    lib.transpose(lib.i(g_off), lib.i(g_idx), n, lib.i(t_off), lib.i(t_idx))
    lib.merge_rows(lib.i(g_off), lib.i(g_idx), lib.i(t_off), lib.i(t_idx), n, lib.i(dummy),
                   lib.i(dummy), lib.i(counts), 0)  # fmt: skip
    off = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    idx = np.empty(max(int(off[-1]), 1), np.int64)
    lib.merge_rows(lib.i(g_off), lib.i(g_idx), lib.i(t_off), lib.i(t_idx), n, lib.i(off),
                   lib.i(idx), lib.i(counts), 1)  # fmt: skip
    return Neighbours(off, idx[: off[-1]].copy(), box_arr)


def brute_force_neighbours(pos, h, box=None) -> list:
    """neighbours.py:74-82: sorted {j : |r_ij| < 2 max(h_i, h_j)} for each i (NumPy, O(N^2)).
    The distance is np.linalg.norm of the minimum-image separation, an independent
    evaluation of the same predicate."""
    pos = np.asarray(pos, float)
    h = np.asarray(h, float)
    box_arr = None if box is None else np.broadcast_to(np.asarray(box, float), (3,))
    out = []
    for i in range(len(pos)):
        dx = minimum_image(pos - pos[i], box_arr)
        r = np.linalg.norm(dx, axis=1)
        out.append(np.flatnonzero(r < R_KERN * np.maximum(h[i], h)))
    return out


def tighten(nb: Neighbours, pos, h, rows=None) -> Neighbours:
    """neighbours.py:306-323: keep pairs with r < 2 max(h_i, h_j); rows not in ``rows`` empty."""
    pos = lib.f64(pos)
    h = lib.f64(h)
    box = np.zeros(3) if nb.box is None else lib.f64(nb.box)
    n = len(pos)
    rows = np.arange(n, dtype=np.int64) if rows is None else lib.i64(rows)
    counts = np.zeros(n, np.int64)
    dummy = np.empty(1, np.int64)
    lib.tighten(lib.d(pos), lib.d(h), lib.d(box), lib.i(nb.offset), lib.i(nb.index), lib.i(rows),
                len(rows), lib.i(dummy), lib.i(dummy), lib.i(counts), 0)  # fmt: skip
    off = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    idx = np.empty(max(int(off[-1]), 1), np.int64)
    lib.tighten(lib.d(pos), lib.d(h), lib.d(box), lib.i(nb.offset), lib.i(nb.index), lib.i(rows),
                len(rows), lib.i(off), lib.i(idx), lib.i(counts), 1)  # fmt: skip
    return Neighbours(off, idx[: off[-1]].copy(), nb.box)


def grid_order(x, h_search):
    """Particle order of the GPU implementation's neighbour grid (its layout policy; the
    reference keeps the caller's order): cubic cells of the fixed-point mean search length (each
    h quantised to h / h_max * 2^32, summed exactly as integers), at least 1/2048 of the widest
    span, at most 2^27 cells (the cell grows by 1.25 until it fits); particles sorted by raster
    cell id, ascending index inside a cell. Open boxes only."""
    x = lib.f64(x)
    h = lib.f64(h_search)
    n = len(x)
    lo = x.min(axis=0)
    span = x.max(axis=0) - lo
    hmax = float(h.max())
    scale = 4294967296.0 / hmax if hmax > 0.0 else 0.0
    hs = int(np.sum((h * scale).astype(np.uint64), dtype=np.uint64))
    hmean = (float(hs) / scale) / float(n) if scale > 0.0 else 0.0
    cell = max(max(hmean, float(np.max(span)) / 2048.0), 1e-300)
    while True:
        nc = [int(max(np.floor(span[c] / cell) + 1.0, 1.0)) for c in range(3)]
        total = 1.0
        for c in range(3):
            total *= float(nc[c])
        if total <= 134217728.0:
            break
        cell *= 1.25
    inv = 1.0 / cell
    f = np.floor((x - lo) * inv)
    c = np.clip(f, 0, np.array(nc) - 1).astype(np.int64)
    key = (c[:, 0] * nc[1] + c[:, 1]) * nc[2] + c[:, 2]
    return np.argsort(key, kind="stable")
