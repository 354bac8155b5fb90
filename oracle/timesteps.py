"""Power-of-two time-step levels on an integer timeline, oracle side (test infrastructure only).

Restates timesteps.py:21-88 of the reference in NumPy: step length of a level in ticks, the level
of a wanted step, the sync test, the rule for moving to a longer step, and the neighbour limiter.
"""

import numpy as np

K_MAX = 40  # timesteps.py:15


def ticks(level):
    """timesteps.py:21-23: 2^(K_MAX - k) ticks."""
    return np.left_shift(np.int64(1), (K_MAX - np.asarray(level, np.int64)))


def level_for_dt(dt_wanted, dt_max):
    """timesteps.py:26-39: smallest level k >= 0 with dt_max / 2^k <= dt_wanted."""
    dt_wanted = np.asarray(dt_wanted, float)
    if np.any(~(dt_wanted > 0)):
        raise ValueError("time steps must be positive")
    k = np.ceil(np.log2(dt_max / dt_wanted) - 1e-12).astype(np.int64)
    k = np.maximum(k, 0)
    k = k + (dt_max / np.exp2(k) > dt_wanted)
    if np.any(k > K_MAX):
        raise ValueError(f"time step needs more than K_MAX = {K_MAX} halvings of dt_max")
    return k


def is_active(n, level):
    """timesteps.py:42-44"""
    return np.mod(np.int64(n), ticks(level)) == 0


def allowed_level(n, current, wanted):
    """timesteps.py:47-60: a shorter step (higher level) always; a longer one only to the longest
    level in [wanted, current) that n is a multiple of."""
    current = np.asarray(current, np.int64)
    wanted = np.asarray(wanted, np.int64)
    new = np.where(wanted >= current, wanted, current)
    for i in np.flatnonzero(wanted < current):
        for k in range(int(wanted[i]), int(current[i])):
            if np.int64(n) % ticks(k) == 0:
                new[i] = k
                break
    return new


def limit_rows(level, rows, offset, index, max_gap):
    """blocks.py:338-353 (and timesteps.py:63-80 with every row): raise the rows' levels until
    each is within max_gap of all its neighbours'. In place."""
    changed = True
    while changed:
        changed = False
        for i in rows:
            nbrs = index[offset[i] : offset[i + 1]]
            need = max(int(level[i]), int(level[nbrs].max()) - max_gap) if len(nbrs) else int(level[i])
            if need > level[i]:
                level[i] = need
                changed = True


def limit_levels(levels, neighbours, max_gap=2):
    """timesteps.py:83-88"""
    out = np.array(levels, np.int64)
    limit_rows(out, np.arange(len(out)), neighbours.offset, neighbours.index, int(max_gap))
    return out
