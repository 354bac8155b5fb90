"""The no-allocation entry point nb_dt_minima_dev_static against nb_dt_minima_dev (GPU).

Both compute the stable_dt minima (coupled.py:213-226) into device memory; the static variant
keeps its partial minima in a fixed device array instead of the context's scratch arena. Minima
are exact, so the two must agree bit for bit: on two calls in a row, the second with a smaller
input (the partial rows of the first call are still in the array), with and without a viscosity
array, and with a NaN that both must propagate. After the arena is released (nb_ctx_trim), the
static variant leaves it empty: it allocates nothing.
"""

import ctypes

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _inputs(rng, n):
    import torch

    dev = torch.device("cuda", 0)
    h = torch.as_tensor(rng.uniform(0.5, 2.0, n), device=dev)
    vsig = torch.as_tensor(rng.uniform(1e-3, 2e-3, n), device=dev)
    acc = torch.as_tensor(rng.normal(0, 1e-6, (n, 3)), device=dev)
    acc[::7] = 0.0  # rows left out of the force criterion
    nu = torch.as_tensor(np.where(rng.uniform(size=n) < 0.5, rng.uniform(1e-4, 1e-3, n), 0.0), device=dev)
    x = torch.as_tensor(rng.normal(0, 50, (n, 3)) + np.array([3000.0, 0.0, 0.0]), device=dev)
    return h, vsig, acc, nu, x


def _both(h, vsig, acc, nu, x):
    import torch

    from nebula import _hip

    n = int(h.shape[0])
    out_a = torch.full((4,), -1.0, dtype=torch.float64, device=h.device)
    out_b = torch.full((4,), -2.0, dtype=torch.float64, device=h.device)
    args = (_hip.ptr(h), _hip.ptr(vsig), _hip.ptr(acc), _hip.ptr(nu), 0.25, _hip.ptr(x), n)
    _hip.call("nb_dt_minima_dev", *args, _hip.ptr(out_a))
    _hip.call("nb_dt_minima_dev_static", *args, _hip.ptr(out_b))
    return out_a.cpu().numpy(), out_b.cpu().numpy()


def test_static_minima_equal_the_arena_variant(rng):
    print("no-alloc: nb_dt_minima_dev_static vs nb_dt_minima_dev, 1048576 then 3000 particles", flush=True)
    for n in (1_048_576, 3000, 3000, 1):
        for with_nu in (True, False):
            h, vsig, acc, nu, x = _inputs(rng, n)
            a, b = _both(h, vsig, acc, nu if with_nu else None, x)
            bit_equal(f"n {n}, nu {with_nu}: minima", a, b)
            check(f"n {n}: finite Courant minimum", True, bool(np.isfinite(b[0])))


def test_static_minima_propagate_nan(rng):
    import torch

    h, vsig, acc, nu, x = _inputs(rng, 5000)
    h[1234] = float("nan")
    a, b = _both(h, vsig, acc, nu, x)
    check("a NaN h gives a NaN Courant minimum (arena variant)", True, bool(np.isnan(a[0])))
    bit_equal("NaN case: minima", a, b)
    torch.cuda.synchronize()


def test_static_variant_allocates_nothing(rng):
    from nebula import _hip

    h, vsig, acc, nu, x = _inputs(rng, 200_000)
    _hip.trim()
    lib = _hip.lib()
    before = int(lib.nb_ctx_scratch_bytes(_hip.ctx()))
    g0 = _hip.generation()
    import torch

    out = torch.empty(4, dtype=torch.float64, device=h.device)
    for _ in range(3):
        _hip.call("nb_dt_minima_dev_static", _hip.ptr(h), _hip.ptr(vsig), _hip.ptr(acc), _hip.ptr(nu), 0.25,
                  _hip.ptr(x), int(h.shape[0]), _hip.ptr(out))  # fmt: skip
    torch.cuda.synchronize()
    check("arena empty after trim", 0, before)
    check("the static variant left the arena empty", 0, int(lib.nb_ctx_scratch_bytes(_hip.ctx())))
    check("and moved nothing", g0, _hip.generation())
    a, b = _both(h, vsig, acc, nu, x)
    bit_equal("after the arena was released: minima", a, b)


def test_static_variant_rejects_bad_sizes():
    from nebula import _hip

    lib = _hip.lib()
    out = ctypes.c_void_p(0)
    rc = lib.nb_dt_minima_dev_static(_hip.ctx(), None, None, None, None, None, 0.25, None, 0, out)
    check("n = 0 and null pointers refused", True, rc < 0)
    check("with a message", True, b"bad arguments" in lib.nb_last_error(_hip.ctx()))
