"""Host logic of the device list layouts (no GPU): RowLists (rows of their own capacity, the
mixed step's layout, rows unsorted) and BlockedLists (64-row blocks, of one capacity or with block
offsets, the FP64 step's layout) convert to the reference's CSR Neighbours row for row."""

import numpy as np


def _rows(rng, n):
    ln = rng.integers(0, 9, n).astype(np.int32)
    ln[5] = 0  # an empty row
    return ln, [rng.choice(n, int(k), replace=False).astype(np.int32) for k in ln]


def test_row_lists_to_csr(rng):
    import torch

    from nebula.sim import neighbours as nbm

    n = 300
    ln, vals = _rows(rng, n)
    cap = ln + rng.integers(0, 3, n)  # padded rows
    start = np.zeros(n + 1, np.int64)
    np.cumsum(cap, out=start[1:])
    idx = np.full(start[-1], -1, np.int32)
    for i in range(n):
        idx[start[i] : start[i] + ln[i]] = vals[i]
    lists = nbm.RowLists(torch.as_tensor(ln), torch.as_tensor(idx), torch.as_tensor(start), int(ln.sum()))
    csr = lists.to_csr()
    assert np.array_equal(csr.offset[1:], np.cumsum(ln))
    assert np.array_equal(csr.index, np.concatenate([np.sort(v) for v in vals]).astype(np.int64))


def test_blocked_lists_to_csr(rng):
    import torch

    from nebula.sim import neighbours as nbm

    n = 300
    nblk = (n + 63) // 64
    ln, vals = _rows(rng, n)
    want = np.concatenate(vals).astype(np.int64)
    # blocks with capacities of their own (block starts)
    cap = [int(ln[b * 64 : (b + 1) * 64].max()) for b in range(nblk)]
    boff = np.zeros(nblk + 1, np.int64)
    np.cumsum(np.array(cap) * 64, out=boff[1:])
    idx = np.full(boff[-1], -1, np.int32)
    for i in range(n):
        for k in range(ln[i]):
            idx[boff[i // 64] + i % 64 + 64 * k] = vals[i][k]
    lists = nbm.BlockedLists(torch.as_tensor(ln), torch.as_tensor(idx), 0, int(ln.sum()), None,
                             block_offset=torch.as_tensor(boff))  # fmt: skip
    assert np.array_equal(lists.to_csr().index, want)
    assert lists.desc().offset is not None
    # one capacity for every block
    ucap = int(ln.max())
    uidx = np.full(nblk * ucap * 64, -1, np.int32)
    for i in range(n):
        for k in range(ln[i]):
            uidx[((i // 64) * ucap + k) * 64 + i % 64] = vals[i][k]
    ulists = nbm.BlockedLists(torch.as_tensor(ln), torch.as_tensor(uidx), ucap, int(ln.sum()), None)
    assert np.array_equal(ulists.to_csr().index, want)
    assert ulists.desc().offset is None
