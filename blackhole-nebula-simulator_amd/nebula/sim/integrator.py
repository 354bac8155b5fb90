"""Generalized leapfrog for the GN Hamiltonian on the GPU (integrator.py of the reference).

``glf_step_many(x, p, dt)`` keeps the reference's contract (integrator.py:188-198): one step of
size dt (scalar or per particle) for N particles, returning (x1, p1, iterations) and raising
GLFConvergenceError naming the first failing particle. The kernel is csrc/step.hip k_glf64.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _hip

TOL = 1e-15
MAX_ITER = 10
R_MIN = 2.5
_FAIL_P_HALF = -1
_FAIL_X1 = -2
_FAIL_RADIUS = -3
_MESSAGES = {
    _FAIL_P_HALF: "stage 1 (p_half) did not converge in 10 iterations: time step too large",
    _FAIL_X1: "stage 2 (x1) did not converge in 10 iterations: time step too large",
    _FAIL_RADIUS: f"particle inside r = {R_MIN}",
}


class GLFConvergenceError(RuntimeError):
    """An implicit stage of the generalized leapfrog failed, or the particle is too deep."""


def glf_device(x_d, p_d, dt, x_out=None, p_out=None, raise_on_failure=True):
    """Device form: x_d, p_d (N,3) float64 tensors; dt a float or (N,) tensor."""
    n = int(x_d.shape[0])
    x_out = _dev.empty((n, 3)) if x_out is None else x_out
    p_out = _dev.empty((n, 3)) if p_out is None else p_out
    iters = _dev.empty(n, torch.int32)
    nfail = torch.zeros(1, dtype=torch.int32, device=x_d.device)
    if isinstance(dt, (float, int, np.floating)):
        dt_arr, dt_s = None, float(dt)
    else:
        dt_arr, dt_s = _dev.f64(dt), 0.0
    _hip.call("nb_glf_step", _hip.ptr(x_d), _hip.ptr(p_d), _hip.ptr(dt_arr), dt_s, n, _hip.ptr(x_out),
              _hip.ptr(p_out), _hip.ptr(iters), _hip.ptr(nfail))  # fmt: skip
    if raise_on_failure and int(nfail.item()) > 0:
        _raise_on_failure(_dev.host(iters))
    return x_out, p_out, iters


def _raise_on_failure(iters: np.ndarray) -> None:
    bad = iters < 0
    if np.any(bad):
        first = int(np.argmax(bad))
        raise GLFConvergenceError(f"particle {first}: {_MESSAGES[int(iters[first])]}")


def glf_step_many(x, p, dt):
    """``glf_step`` for N particles at once. x, p: (N, 3) float64; dt a scalar or (N,).
    Returns (x1, p1, iterations per particle). Raises GLFConvergenceError."""
    as_dev = _dev.is_dev(x)
    x_d = _dev.f64(x)
    if np.ndim(dt) == 0 and not isinstance(dt, torch.Tensor):
        dt_v = float(dt)
    else:
        dt_v = _dev.f64(np.broadcast_to(np.asarray(dt, np.float64), (len(x_d),)) if not isinstance(dt, torch.Tensor)
                        else dt)  # fmt: skip
    x1, p1, it = glf_device(x_d, _dev.f64(p), dt_v)
    if as_dev:
        return x1, p1, it
    return _dev.host(x1), _dev.host(p1), _dev.host(it)
