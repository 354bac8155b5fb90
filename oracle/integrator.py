"""GN gravity, oracle side (test infrastructure only): generalized leapfrog
(integrator.py:139-198) and v = dH/dp (fastphys.py:24-42)."""

import numpy as np

from . import lib

MESSAGES = {
    -1: "stage 1 (p_half) did not converge in 10 iterations: time step too large",
    -2: "stage 2 (x1) did not converge in 10 iterations: time step too large",
    -3: "particle inside r = 2.5",
}


class GLFConvergenceError(RuntimeError):
    pass


def glf_step_many(x, p, dt):
    """integrator.py:188-198: (x1, p1, iterations); raises on a negative code."""
    x = lib.f64(x)
    p = lib.f64(p)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(x),)))
    x_out, p_out = np.empty_like(x), np.empty_like(p)
    iters = np.empty(len(x), np.int64)
    lib.glf_step_many(lib.d(x), lib.d(p), lib.d(dt), len(x), lib.d(x_out), lib.d(p_out),
                      lib.i(iters))  # fmt: skip
    bad = iters < 0
    if np.any(bad):
        first = int(np.argmax(bad))
        raise GLFConvergenceError(f"particle {first}: {MESSAGES[int(iters[first])]}")
    return x_out, p_out, iters


def glf_step_many_codes(x, p, dt):
    """As glf_step_many, but returns the failure codes instead of raising."""
    x = lib.f64(x)
    p = lib.f64(p)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(x),)))
    x_out, p_out = np.empty_like(x), np.empty_like(p)
    iters = np.empty(len(x), np.int64)
    lib.glf_step_many(lib.d(x), lib.d(p), lib.d(dt), len(x), lib.d(x_out), lib.d(p_out),
                      lib.i(iters))  # fmt: skip
    return x_out, p_out, iters


def gn_v_from_p(x, p):
    """fastphys.py:35-42"""
    x = lib.f64(x)
    p = lib.f64(p)
    out = np.empty_like(p)
    lib.gn_v_from_p(lib.d(x), lib.d(p), len(x), lib.d(out))
    return out
