"""Neighbour lists on the GPU, behind the reference's call surface (neighbours.py).

``find_neighbours_numba(pos, h, box)`` and ``tighten(nb, pos, h, rows)`` return the same
symmetric, sorted CSR lists as the reference (neighbours.py:240-323): row i lists every j with
|x_i - x_j| < 2 max(h_i, h_j), itself included. The lists are built by the HIP cell-grid
search (csrc/lists.hip) and stay on the device; ``offset`` / ``index`` give NumPy int64 copies
on first access, as the reference's dataclass fields are.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _hip

R_KERN = 2.0


class Neighbours:
    """CSR neighbour lists (neighbours.py:22-36): the neighbours of i are
    index[offset[i]:offset[i + 1]] (sorted, including i). ``box``: periodic lengths or None.
    Device copies: ``d_offset`` (int64) and ``d_index`` (int32)."""

    def __init__(self, offset=None, index=None, box=None, *, d_offset=None, d_index=None):
        self._offset = None if offset is None else np.asarray(offset, np.int64)
        self._index = None if index is None else np.asarray(index, np.int64)
        self.box = None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()
        self._d_offset = d_offset
        self._d_index = d_index

    @classmethod
    def from_device(cls, d_offset: torch.Tensor, d_index: torch.Tensor, box) -> "Neighbours":
        return cls(None, None, box, d_offset=d_offset, d_index=d_index)

    @property
    def offset(self) -> np.ndarray:
        if self._offset is None:
            self._offset = _dev.host(self._d_offset)
        return self._offset

    @property
    def index(self) -> np.ndarray:
        if self._index is None:
            # the reference's index array is int64 (neighbours.py:22-36); the device keeps int32
            self._index = _dev.host(self._d_index).astype(np.int64)
        return self._index

    @property
    def d_offset(self) -> torch.Tensor:
        if self._d_offset is None:
            self._d_offset = _dev.i64(self._offset)
        return self._d_offset

    @property
    def d_index(self) -> torch.Tensor:
        if self._d_index is None:
            idx = self._index
            if idx.size and (idx.max() >= 2**31 - 1 or idx.min() < 0):
                raise ValueError("neighbour indices out of range")
            self._d_index = _dev.i32(idx)
        return self._d_index

    @property
    def n_entries(self) -> int:
        if self._d_offset is not None:
            return int(self._d_offset[-1])
        return int(self._offset[-1])

    def of(self, i: int) -> np.ndarray:
        return self.index[self.offset[i] : self.offset[i + 1]]

    @property
    def counts(self) -> np.ndarray:
        return np.diff(self.offset)

    def desc(self) -> _hip.Lists:
        """nb_lists descriptor (CSR); the tensors stay referenced by this object."""
        d = _hip.Lists()
        d.offset = self.d_offset.data_ptr()
        d.index = self.d_index.data_ptr()
        d.length = None
        d.row_cap = 0
        d.n_entries = self.n_entries
        d.n_rows = int(self.d_offset.numel()) - 1
        return d


class BlockedLists:
    """Neighbour lists in 64-row blocks on the device (the FP64 step's layout): row i has
    ``length[i]`` entries, entry k at ((i // 64) * row_cap + k) * 64 + i % 64, or, when
    ``block_offset`` (N // 64 + 1 int64 block starts) is given, at block_offset[i // 64] +
    64 k + i % 64 (blocks with capacities of their own)."""

    def __init__(self, length: torch.Tensor, index: torch.Tensor, row_cap: int, n_total: int, box=None,
                 block_offset: torch.Tensor | None = None):  # fmt: skip
        self.length = length
        self.index = index
        self.row_cap = int(row_cap)
        self.n_total = int(n_total)
        self.box = box
        self.block_offset = block_offset

    @property
    def n_entries(self) -> int:
        return int(self.index.numel())

    def desc(self) -> _hip.Lists:
        d = _hip.Lists()
        d.offset = None if self.block_offset is None else self.block_offset.data_ptr()
        d.index = self.index.data_ptr()
        d.length = self.length.data_ptr()
        d.row_cap = self.row_cap
        d.n_entries = self.n_entries
        d.n_rows = int(self.length.numel())
        return d

    def to_csr(self) -> Neighbours:
        """The same lists as a CSR Neighbours (host arrays)."""
        n = int(self.length.numel())
        ln = self.length.cpu().numpy().astype(np.int64)
        idx = self.index.cpu().numpy()
        if self.block_offset is None:
            start = np.arange(0, (n + 63) // 64, dtype=np.int64) * self.row_cap * 64
        else:
            start = self.block_offset.cpu().numpy()
        off = np.zeros(n + 1, np.int64)
        np.cumsum(ln, out=off[1:])
        row = np.repeat(np.arange(n, dtype=np.int64), ln)
        k = np.arange(off[-1], dtype=np.int64) - off[row]  # entry k of its row
        out = idx[start[row // 64] + row % 64 + 64 * k].astype(np.int64)
        return Neighbours(off, out, self.box)


class RowLists:
    """Neighbour lists in rows of their own capacity (the mixed-precision step's layout): row i
    has ``length[i]`` entries at index[offset[i] : offset[i] + length[i]], not sorted. ``offset``
    has N + 1 entries."""

    def __init__(self, length: torch.Tensor, index: torch.Tensor, offset: torch.Tensor, n_total: int, box=None):
        self.length = length
        self.index = index
        self.offset = offset
        self.n_total = int(n_total)
        self.box = box

    @property
    def n_entries(self) -> int:
        return int(self.index.numel())

    def desc(self) -> _hip.Lists:
        d = _hip.Lists()
        d.offset = self.offset.data_ptr()
        d.index = self.index.data_ptr()
        d.length = self.length.data_ptr()
        d.row_cap = 0
        d.n_entries = self.n_entries
        d.row_major = 1
        d.n_rows = int(self.length.numel())
        return d

    def to_csr(self) -> Neighbours:
        """The same lists as a CSR Neighbours (host arrays, rows sorted)."""
        n = int(self.length.numel())
        ln = self.length.cpu().numpy().astype(np.int64)
        start = self.offset.cpu().numpy()
        idx = self.index.cpu().numpy()
        off = np.zeros(n + 1, np.int64)
        np.cumsum(ln, out=off[1:])
        row = np.repeat(np.arange(n, dtype=np.int64), ln)
        k = np.arange(off[-1], dtype=np.int64) - off[row]
        vals = idx[start[row] + k].astype(np.int64)
        out = np.sort(row * max(n, 1) + vals) - row * max(n, 1)  # each row sorted in place
        return Neighbours(off, out, self.box)


class IndexBuffer:
    """Reusable device storage for the tight lists of a time-stepping loop. The list size drifts
    from step to step; a fresh allocation per step would leave the caching allocator with ever
    larger blocks split by the step's smaller tensors. take() hands out a view of one buffer that
    grows by a quarter beyond the request when it is too small. The lists of a step are only valid
    until the next take()."""

    def __init__(self):
        self._t = None

    def take(self, numel: int) -> torch.Tensor:
        numel = max(int(numel), 1)
        if self._t is None or self._t.numel() < numel:
            self._t = None
            self._t = _dev.empty(numel + numel // 4 + 4096, torch.int32)
        return self._t[:numel]


def _blocked_alloc(n: int, row_cap: int, index_buffer: IndexBuffer | None = None):
    nblk = (n + 63) // 64
    size = max(nblk * row_cap * 64, 1)
    idx = _dev.empty(size, torch.int32) if index_buffer is None else index_buffer.take(size)
    return _dev.empty(n, torch.int32), idx


def build_lists_blocked(pos_d: torch.Tensor, h_d: torch.Tensor, box_arr, row_cap_hint: int = 0) -> BlockedLists:
    """Symmetric lists of device positions at search lengths h, in 64-row blocks."""
    n = int(pos_d.shape[0])
    cap = max(int(row_cap_hint), 16)
    need = np.zeros(1, np.int64)
    total = np.zeros(1, np.int64)
    ln, idx = _blocked_alloc(n, cap)
    rc = _hip.call("nb_lists_build_blocked", _hip.ptr(pos_d), _hip.ptr(h_d), n, _hip.box_arg(box_arr), _hip.ptr(ln),
                   _hip.ptr(idx), cap, need.ctypes.data_as(_hip._P64), total.ctypes.data_as(_hip._P64))  # fmt: skip
    if rc == _hip.NB_RETRY_CAPACITY:
        cap = int(need[0]) + int(need[0]) // 8 + 8
        ln, idx = _blocked_alloc(n, cap)
        _hip.call("nb_lists_emit_blocked", _hip.ptr(ln), _hip.ptr(idx), cap)
    return BlockedLists(ln, idx, cap, int(total[0]), box_arr)


def tighten_blocked(nb: BlockedLists, pos_d: torch.Tensor, h_d: torch.Tensor) -> BlockedLists:
    """tighten (neighbours.py:306-323) of blocked lists into blocked lists of the same capacity."""
    n = int(pos_d.shape[0])
    ln, idx = _blocked_alloc(n, nb.row_cap)
    total = np.zeros(1, np.int64)
    _hip.call("nb_tighten_blocked", _hip.ptr(pos_d), _hip.ptr(h_d), n, _hip.box_arg(nb.box), _hip.ptr(nb.length),
              _hip.ptr(nb.index), nb.row_cap, _hip.ptr(ln), _hip.ptr(idx), total.ctypes.data_as(_hip._P64))  # fmt: skip
    return BlockedLists(ln, idx, nb.row_cap, int(total[0]), nb.box)


def grid_order(pos_d: torch.Tensor, h_d: torch.Tensor, box_arr=None, cell_factor: float = 1.0) -> torch.Tensor:
    """The neighbour grid's particle order (int32) for these positions and search lengths."""
    n = int(pos_d.shape[0])
    order = _dev.empty(n, torch.int32)
    _hip.call("nb_grid_order", _hip.ptr(pos_d), _hip.ptr(h_d), n, _hip.box_arg(box_arr), float(cell_factor),
              _hip.ptr(order))  # fmt: skip
    return order


def grid_order_seg(pos_d: torch.Tensor, h_d: torch.Tensor, cloud_d: torch.Tensor, seg_off: np.ndarray,
                   cell_factor: float = 1.0) -> torch.Tensor:  # fmt: skip
    """grid_order for an ensemble of independent systems (seg_off: host int64 ranges)."""
    n = int(pos_d.shape[0])
    order = _dev.empty(n, torch.int32)
    off = np.ascontiguousarray(seg_off, np.int64)
    _hip.call("nb_grid_order_seg", _hip.ptr(pos_d), _hip.ptr(h_d), n, _hip.ptr(cloud_d),
              off.ctypes.data_as(_hip._P64), len(off) - 1, float(cell_factor), _hip.ptr(order))  # fmt: skip
    return order


def permute(t: torch.Tensor, order: torch.Tensor) -> torch.Tensor:
    """Rows of t (float64 (N,...) or int64 (N,)) taken in ``order``."""
    n = int(order.numel())
    out = torch.empty_like(t)
    if t.dtype == torch.int64:
        _hip.call("nb_permute_i64", _hip.ptr(t), _hip.ptr(order), n, _hip.ptr(out))
    else:
        width = int(t.numel() // max(n, 1))
        _hip.call("nb_permute", _hip.ptr(t), _hip.ptr(order), n, width, _hip.ptr(out))
    return out


def permute_many(tensors, order: torch.Tensor, kicks=None, s_dev=None, scale: float = 0.5) -> list:
    """permute() of several float64 / int64 tensors with one kernel launch. kicks: {position:
    (y, a)} of float64 pairs whose moved kick y + s a (s = scale * s_dev[0], read on the device)
    takes that position of the result (its entry in ``tensors`` is None)."""
    import ctypes

    n = int(order.numel())
    kicks = kicks or {}
    srcs = [kicks[k][0] if k in kicks else t for k, t in enumerate(tensors)]
    for t in srcs:
        if t.dtype not in (torch.float64, torch.int64) or not t.is_contiguous() or t.shape[0] != n:
            raise ValueError("permute_many takes contiguous float64 / int64 tensors of one length")
    for y, a in kicks.values():
        if y.dtype != torch.float64 or a.dtype != torch.float64 or a.shape != y.shape or not a.is_contiguous():
            raise ValueError("a kick pairs two contiguous float64 tensors of one shape")
    outs = [torch.empty_like(t) for t in srcs]
    k = len(srcs)
    src = (ctypes.c_void_p * k)(*[t.data_ptr() for t in srcs])
    dst = (ctypes.c_void_p * k)(*[o.data_ptr() for o in outs])
    width = (ctypes.c_int32 * k)(*[int(t.numel() // max(n, 1)) for t in srcs])
    if not kicks:
        _hip.call("nb_permute_many", k, src, dst, width, _hip.ptr(order), n)
        return outs
    if s_dev is None:
        raise ValueError("kicks in the permutation take their step from the device (s_dev)")
    src2 = (ctypes.c_void_p * k)(*[kicks[j][1].data_ptr() if j in kicks else None for j in range(k)])
    _hip.call("nb_permute_kick_many", k, src, src2, dst, width, _hip.ptr(s_dev), float(scale), _hip.ptr(order), n)
    return outs


def minimum_image(dx: np.ndarray, box) -> np.ndarray:
    """Separation vectors folded into (-L/2, L/2] (neighbours.py:39-43)."""
    if box is None:
        return dx
    return dx - box * np.round(dx / box)


def wrap(pos: np.ndarray, box) -> np.ndarray:
    """Positions folded into [0, L) (neighbours.py:46-50)."""
    if box is None:
        return pos
    return np.mod(pos, box)


def _box3(box):
    return None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()


def build_lists(pos_d: torch.Tensor, h_d: torch.Tensor, box_arr, capacity_hint: int = 0) -> Neighbours:
    """Symmetric lists of device positions (N,3) at search lengths h (N,), kept on device."""
    n = int(pos_d.shape[0])
    off = _dev.empty(n + 1, torch.int64)
    total = np.zeros(1, np.int64)
    cap = max(int(capacity_hint), 1)
    idx = _dev.empty(cap, torch.int32)
    rc = _hip.call("nb_lists_build", _hip.ptr(pos_d), _hip.ptr(h_d), n, _hip.box_arg(box_arr),
                   _hip.ptr(off), _hip.ptr(idx), cap, total.ctypes.data_as(_hip._P64))  # fmt: skip
    if rc == _hip.NB_RETRY_CAPACITY:
        cap = int(total[0])
        idx = _dev.empty(max(cap, 1), torch.int32)
        _hip.call("nb_lists_emit", _hip.ptr(idx), max(cap, 1))
    m = int(total[0])
    return Neighbours.from_device(off, idx[:m], box_arr)


def find_neighbours_numba(pos, h, box=None) -> Neighbours:
    """Symmetric neighbour lists for positions (N, 3) and smoothing lengths (N,)
    (neighbours.py:240-281). Raises ValueError if 2 max(h) reaches half a periodic box."""
    box_arr = _box3(box)
    h_np = h.detach().cpu().numpy() if isinstance(h, torch.Tensor) else np.asarray(h, np.float64)
    if box_arr is not None and len(h_np) and np.any(R_KERN * h_np.max() >= 0.5 * box_arr):
        raise ValueError("kernel support exceeds half the periodic box: neighbours ambiguous")
    return build_lists(_dev.f64(pos), _dev.f64(h), box_arr)


# search lengths of find_neighbours' first pass: a few ulps above h, so that every pair at
# exactly 2 max(h_i, h_j) is inside the strict search
_INCLUSIVE_MARGIN = 1.0 + 8.0 * np.finfo(np.float64).eps


def find_neighbours(pos, h, box=None, workers: int = -1) -> Neighbours:
    """The reference's cKDTree twin (neighbours.py:53-71): symmetric lists of every j with
    |x_i - x_j| <= 2 max(h_i, h_j), the boundary included as query_ball_point includes it
    (find_neighbours_numba's grid search is strict). Lists are built at search lengths a few ulps
    above h, then tightened with the inclusive test at h. ``workers`` is accepted for signature
    compatibility (the GPU decides its own parallelism)."""
    del workers
    box_arr = _box3(box)
    h_np = h.detach().cpu().numpy() if isinstance(h, torch.Tensor) else np.asarray(h, np.float64)
    if box_arr is not None and len(h_np) and np.any(R_KERN * h_np.max() >= 0.5 * box_arr):
        raise ValueError("kernel support exceeds half the periodic box: neighbours ambiguous")
    pos_d, h_d = _dev.f64(pos), _dev.f64(h)
    wide = build_lists(pos_d, h_d * _INCLUSIVE_MARGIN, box_arr)
    return tighten(wide, pos_d, h_d, inclusive=True)


def tighten(nb: Neighbours, pos, h, rows=None, inclusive: bool = False) -> Neighbours:
    """Keep only the pairs with r < 2 max(h_i, h_j) (neighbours.py:306-323). ``rows``: lists
    only for these particles, the others' left empty. ``inclusive`` keeps r <= 2 max(h_i, h_j)
    (find_neighbours' pair set)."""
    pos_d = _dev.f64(pos)
    h_d = _dev.f64(h)
    n = int(pos_d.shape[0])
    rows_d, nrows = _dev.rows_dev(rows, n)
    cap = max(nb.n_entries, 1)
    off = _dev.empty(n + 1, torch.int64)
    idx = _dev.empty(cap, torch.int32)
    total = np.zeros(1, np.int64)
    _hip.call("nb_tighten_inclusive" if inclusive else "nb_tighten", _hip.ptr(pos_d), _hip.ptr(h_d), n,
              _hip.box_arg(nb.box), _hip.ptr(nb.d_offset),
              _hip.ptr(nb.d_index), _hip.ptr(rows_d), nrows, _hip.ptr(off), _hip.ptr(idx), cap,
              total.ctypes.data_as(_hip._P64))  # fmt: skip
    return Neighbours.from_device(off, idx[: int(total[0])], nb.box)
