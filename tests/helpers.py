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


def lattice_sphere(n_target, radius, seed=1, jitter=0.1):
    """The n_target points of a jittered cubic lattice nearest its centre, scaled to radius (the
    reference cuts its Wendland glass instead; any near-uniform sphere serves the checks that
    use it)."""
    rng = np.random.default_rng(seed)
    a = (4.0 / 3.0 * np.pi / n_target) ** (1.0 / 3.0)
    k = int(np.ceil(1.0 / a)) + 2
    g = np.arange(-k, k + 1) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = x + rng.uniform(-jitter * a, jitter * a, x.shape)
    x = x[np.argsort(np.linalg.norm(x, axis=1), kind="stable")[:n_target]]
    return x * (radius / np.linalg.norm(x, axis=1).max())


ELL_BLOB = 7.1951756691  # expected-values §2: the r_a = 500 -> r_p = 25 orbit (blob.py:23)
U_BLOB = 4e-8  # blob.py:29
T_END_BLOB = 18_000.0  # blob.py:34


def hot_blob_ic(mass_noise=0.0, seed=2):
    """blob.py:36-52: ~500 particles of radius 3 at apocentre x = (500, 0, 0), momentum
    p = (0, ell / 500, 0), u = U_BLOB, each of mass 1/n (times 1 + mass_noise N(0, 1) when
    mass_noise is set), h = 1.6 (m / rho)^(1/3). Returns x, v, u, mass, h."""
    from oracle.integrator import gn_v_from_p

    pos = lattice_sphere(500, 3.0) + np.array([500.0, 0.0, 0.0])
    n = len(pos)
    p0 = np.tile([0.0, ELL_BLOB / 500.0, 0.0], (n, 1))
    rho = 1.0 / (4 / 3 * np.pi * 27.0)
    mass = np.full(n, 1.0 / n)
    if mass_noise:
        mass = mass * (1.0 + mass_noise * np.random.default_rng(seed).standard_normal(n))
    return pos, gn_v_from_p(pos, p0), np.full(n, U_BLOB), mass, np.full(n, 1.6 * (1 / n / rho) ** (1 / 3))


def hot_blob_config(eta_orb, courant, **kw):
    """blob.py:86-91: the reference's hydro defaults (Wendland C4, Cullen-Dehnen switch,
    reconstruction, conductivity with the velocity signal) at the given step factors."""
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.hydro import HydroConfig
    from nebula.sim.sph import COND_VELOCITY

    hydro_kw = {k: kw.pop(k) for k in list(kw) if k in HydroConfig.__dataclass_fields__}
    return CoupledConfig(hydro=HydroConfig(courant=courant, conductivity_signal=COND_VELOCITY, **hydro_kw),
                         eta_orb=eta_orb, **kw)  # fmt: skip
