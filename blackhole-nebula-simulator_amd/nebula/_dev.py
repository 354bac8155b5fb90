"""Moving arrays between the caller (NumPy or torch) and device tensors."""

from __future__ import annotations

import numpy as np
import torch

from . import _hip


def is_dev(a) -> bool:
    return isinstance(a, torch.Tensor) and a.is_cuda


def f64(a, *, copy: bool = False) -> torch.Tensor:
    """Contiguous float64 tensor on the GPU (no copy if it already is one, unless asked)."""
    dev = _hip.torch_device()
    if isinstance(a, torch.Tensor):
        t = a.to(device=dev, dtype=torch.float64)
        t = t.contiguous()
        return t.clone() if (copy and t.data_ptr() == a.data_ptr()) else t
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def i64(a) -> torch.Tensor:
    dev = _hip.torch_device()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.int64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to(dev)


def i32(a) -> torch.Tensor:
    dev = _hip.torch_device()
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)


def empty(shape, dtype=torch.float64) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=_hip.torch_device())


def host(t) -> np.ndarray:
    """NumPy copy of a tensor (int32 results come back as int64, as the reference's do)."""
    if t is None:
        return None
    a = t.detach().cpu().numpy()
    if a.dtype == np.int32:
        a = a.astype(np.int64)
    return a


def rows_dev(rows, n: int):
    """Device int64 row list, or None for all rows (the kernels then use q itself)."""
    if rows is None:
        return None, n
    r_host = rows.detach().cpu().numpy() if isinstance(rows, torch.Tensor) else np.asarray(rows, np.int64)
    if r_host.size and (r_host.min() < 0 or r_host.max() >= n):
        raise IndexError(f"rows out of range [0, {n})")
    r = i64(rows)
    return r, int(r.numel())


def uniform_mass(mass_dev: torch.Tensor) -> float:
    """The common mass if every particle has the same one, else 0."""
    if mass_dev.numel() == 0:
        return 0.0
    m0 = mass_dev[0]
    if bool(torch.all(mass_dev == m0)) and float(m0) > 0.0:
        return float(m0)
    return 0.0
