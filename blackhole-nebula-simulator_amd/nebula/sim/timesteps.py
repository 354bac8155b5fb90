"""Individual power-of-two time steps on an integer timeline (timesteps.py of the reference).

Time is t = t0 + n tick with n an int64 counter and tick = dt_max / 2^K_MAX. A particle on level
k takes steps of 2^(K_MAX - k) ticks and is at a sync point when n is a multiple of that. The
level arithmetic of a system step runs on the device (nb_block_levels, nb_block_limit,
nb_block_land in csrc/blocks.hip); these functions keep the reference's host-side API.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _hip

K_MAX = 40  # timesteps.py:15


def ticks(level) -> np.ndarray:
    """Step length in ticks of level k: 2^(K_MAX - k) (timesteps.py:21-23)."""
    return np.left_shift(np.int64(1), (K_MAX - np.asarray(level, np.int64)))


def level_for_dt(dt_wanted, dt_max: float) -> np.ndarray:
    """The smallest level k >= 0 with dt_max / 2^k <= dt_wanted (timesteps.py:26-39), found by
    halving dt_max exactly rather than through log2. Raises ValueError for steps that are not
    positive or need more than K_MAX halvings."""
    dt = np.asarray(dt_wanted, float)
    if np.any(~(dt > 0)):
        raise ValueError("time steps must be positive")
    k = np.zeros(dt.shape, np.int64)
    for _ in range(K_MAX + 1):
        over = np.ldexp(float(dt_max), (-k).astype(np.int32)) > dt
        if not np.any(over):
            break
        k = k + over
    if np.any(np.ldexp(float(dt_max), (-k).astype(np.int32)) > dt) or np.any(k > K_MAX):
        raise ValueError(f"time step needs more than K_MAX = {K_MAX} halvings of dt_max")
    return k


def is_active(n: int, level) -> np.ndarray:
    """True where n is a sync point of the level (timesteps.py:42-44)."""
    return np.mod(np.int64(n), ticks(level)) == 0


def allowed_level(n: int, current, wanted) -> np.ndarray:
    """New level at sync point n (timesteps.py:47-60): a shorter step always; a longer one only to
    the longest level in [wanted, current) that n is a multiple of."""
    current = np.asarray(current, np.int64)
    wanted = np.asarray(wanted, np.int64)
    new = np.where(wanted >= current, wanted, current)
    for i in np.flatnonzero(wanted < current):
        for k in range(int(wanted[i]), int(current[i])):
            if np.int64(n) % ticks(k) == 0:
                new[i] = k
                break
    return new


def limit_rows_device(level_d: torch.Tensor, rows_d: torch.Tensor, nb, max_gap: int) -> int:
    """Raise level_d[rows] (int64, in place) until each row is within max_gap of all its
    neighbours (blocks.py:338-353); sweeps of nb_block_limit until none rises. Returns the
    number of sweeps."""
    nrows = int(rows_d.numel())
    if nrows == 0:
        return 0
    changed = torch.zeros(1, dtype=torch.int32, device=level_d.device)
    sweeps = 0
    while True:
        changed.zero_()
        _hip.call("nb_block_limit", _hip.ptr(rows_d), nrows, _hip.ptr(nb.d_offset), _hip.ptr(nb.d_index),
                  int(max_gap), _hip.ptr(level_d), _hip.ptr(changed))  # fmt: skip
        sweeps += 1
        if int(changed.item()) == 0:
            return sweeps


def limit_levels(levels, neighbours, max_gap: int = 2) -> np.ndarray:
    """Neighbour time-step limiter (timesteps.py:83-88): levels of neighbouring particles differ
    by at most max_gap; steps are only shortened. Returns a new array."""
    lv = _dev.i64(np.array(levels, np.int64))
    rows = torch.arange(int(lv.numel()), dtype=torch.int64, device=lv.device)
    limit_rows_device(lv, rows, neighbours, int(max_gap))
    return _dev.host(lv)
