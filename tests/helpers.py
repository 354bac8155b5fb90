"""Comparison helper modelled on the reference's tests/helpers.py:51-96: passes when
|actual - expected| <= atol + rtol |expected| everywhere; per_vector measures 3-vectors by
length; with no tolerance the values must be equal."""

import numpy as np


def check(name, expected, actual, rtol=None, atol=None, per_vector=False):
    numeric = rtol is not None or atol is not None
    if numeric:
        e = np.asarray(expected, dtype=np.float64)
        a = np.asarray(actual, dtype=np.float64)
        if e.shape != a.shape and e.size != 1:
            raise AssertionError(f"{name}: shape mismatch {e.shape} vs {a.shape}")
        if per_vector:
            err = np.linalg.norm(a - e, axis=-1)
            size = np.linalg.norm(np.broadcast_to(e, a.shape), axis=-1)
        else:
            err = np.abs(a - e)
            size = np.abs(np.broadcast_to(e, a.shape))
        limit = (atol or 0.0) + (rtol or 0.0) * size
        ok = bool(np.all(err <= limit)) and not np.any(np.isnan(a))
        worst = float(np.max(err - limit)) if a.size else 0.0
        msg = f"{name}: max abs err {float(err.max()) if err.size else 0:.3g}, worst excess {worst:.3g}"
    else:
        ok = bool(np.all(np.asarray(expected == actual)))
        msg = f"{name}: {expected!r} != {actual!r}" if not ok else name
    print(("PASS " if ok else "FAIL ") + msg)
    assert ok, msg


def bit_equal(name, expected, actual):
    """Bitwise equality (floats compared by their IEEE bit patterns, ints by value)."""
    e = np.ascontiguousarray(expected)
    a = np.ascontiguousarray(actual)
    if e.shape != a.shape:
        raise AssertionError(f"{name}: shape {e.shape} vs {a.shape}")
    if e.dtype.kind in "iub" and a.dtype.kind in "iub":
        diff = e.astype(np.int64) != a.astype(np.int64)
    else:
        diff = e.astype(np.float64).view(np.int64) != a.astype(np.float64).view(np.int64)
    bad = np.flatnonzero(diff.reshape(-1))
    if bad.size:
        k = bad[0]
        print(f"FAIL {name}: {bad.size} of {e.size} differ; first flat index {k}: "
              f"{e.reshape(-1)[k]!r} vs {a.reshape(-1)[k]!r}")
    else:
        print(f"PASS {name} (bit-identical, {e.size} values)")
    assert bad.size == 0, name
