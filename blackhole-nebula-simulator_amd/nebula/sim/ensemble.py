"""Ensembles of independent gas clouds stepped together (no counterpart in the reference, whose
drivers run one cloud per process).

The clouds are stored one after the other in one set of device arrays. A per-particle ``cloud``
index and host ``seg_off`` ranges identify each cloud. Every pass of the step is one launch
over all clouds:

- The neighbour grid is segmented (``nb_grid_order_seg`` / ``nb_hydro_density_mixed`` with
  segments): each cloud gets its own grid planned from its own box and search lengths, and a
  particle searches only its own cloud's cells. Clouds never interact, wherever they are.
- Each cloud keeps its own time step (``stable_dt_clouds``, one segmented reduction). The kicks
  and the GN flow take a step per particle.

Small clouds alone cannot fill the GPU, and the host's per-launch cost dominates their steps.
In one launch, the work of all clouds is spread over all compute units at once.

The physics is the plain method of the cloud workloads (``coupled.baseline_config``) in the
mixed-precision mode, without cooling, the viscosity switch or the learned start.
"""

from __future__ import annotations

import math
from dataclasses import replace

import numpy as np
import torch

from .. import _dev, _hip
from . import coupled
from .coupled import GasState, _hydro, p_from_v, velocity
from .integrator import glf_device
from .neighbours import permute_many


def _check_config(cfg: coupled.CoupledConfig) -> None:
    hc = cfg.hydro
    if hc.precision != "mixed":
        raise ValueError("ensembles run in the mixed-precision mode")
    if math.isfinite(cfg.cooling_beta) or cfg.cooling_powerlaw or hc.switch or hc.learned_h:
        raise ValueError("ensembles run the plain cloud physics (no cooling, switch or learned start)")


def initialise(ics, cfg: coupled.CoupledConfig, cell_factor: float = 1.0) -> GasState:
    """One state holding every cloud of ``ics`` (a list of nebula.ics.IC), p from v."""
    _check_config(cfg)
    sizes = [len(ic.x) for ic in ics]
    off = np.zeros(len(ics) + 1, np.int64)
    np.cumsum(sizes, out=off[1:])
    x = np.concatenate([np.asarray(ic.x, float) for ic in ics])
    v = np.concatenate([np.asarray(ic.v, float) for ic in ics])
    u = np.concatenate([np.asarray(ic.u, float) * np.ones(len(ic.x)) for ic in ics])
    mass = np.concatenate([np.asarray(ic.mass, float) * np.ones(len(ic.x)) for ic in ics])
    h = np.concatenate([np.asarray(ic.h, float) * np.ones(len(ic.x)) for ic in ics])
    cloud = np.repeat(np.arange(len(ics), dtype=np.int32), sizes)
    hc = cfg.hydro
    st = GasState(
        0.0, _dev.f64(x), _dev.f64(p_from_v(cfg, x, v)), _dev.f64(u), _dev.f64(mass), _dev.f64(h),
        _dev.f64(np.full(len(x), hc.alpha_const)),
        e_radiated_dev=torch.zeros(1, dtype=torch.float64, device=_hip.torch_device()), m_uniform=0.0,
        cloud=torch.as_tensor(cloud, device=_hip.torch_device()), seg_off=off, cell_factor=float(cell_factor),
    )  # fmt: skip
    order = coupled.grid_permutation(cfg, st.x, st.h, st)
    x_, p_, u_, m_, h_, a_ = permute_many((st.x, st.p, st.u, st.mass, st.h, st.alpha), order)
    st = replace(st, x=x_, p=p_, u=u_, mass=m_, h=h_, alpha=a_, perm=order.to(torch.int64))
    return _hydro(st, cfg, velocity(cfg, st.x, st.p), 0.0, None)


def stable_dt_clouds(st: GasState, cfg: coupled.CoupledConfig) -> np.ndarray:
    """coupled.stable_dt of every cloud (one segmented reduction)."""
    hc = cfg.hydro
    s = len(st.seg_off) - 1
    out = np.zeros(4 * s)
    _hip.call("nb_dt_minima_seg", _hip.ptr(st.h), _hip.ptr(st.vsig_max), _hip.ptr(st.a_hydro), _hip.ptr(st.nu), 0.25,
              _hip.ptr(st.x), st.seg_off.ctypes.data_as(_hip._P64), s, out.ctypes.data_as(_hip._PD))  # fmt: skip
    m = out.reshape(s, 4)
    dt = hc.courant * m[:, 0]
    dt = np.where(np.isfinite(m[:, 1]), np.minimum(dt, hc.c_force * m[:, 1]), dt)
    if st.nu is not None:
        dt = np.minimum(dt, m[:, 2])
    if cfg.hole:
        # scalar pow per cloud, as coupled.stable_dt takes it: NumPy's vectorised power may differ
        # from the scalar (libm) pow in the last bit on some hosts
        dt = np.minimum(dt, np.array([cfg.eta_orb * float(r) ** 1.5 for r in m[:, 3]]))
    return dt


def step(st: GasState, cfg: coupled.CoupledConfig, dt_clouds: np.ndarray) -> GasState:
    """One step of every cloud with its own dt (coupled.step's split: half kick, GN flow over dt,
    hydro at the new positions with the predicted momenta, half kick)."""
    n = int(st.x.shape[0])
    dt_seg = _dev.f64(np.asarray(dt_clouds, float))
    dt_p, half = _dev.empty(n), _dev.empty(n)
    _hip.call("nb_expand_seg", _hip.ptr(dt_seg), _hip.ptr(st.cloud), n, 1.0, _hip.ptr(dt_p))
    _hip.call("nb_expand_seg", _hip.ptr(dt_seg), _hip.ptr(st.cloud), n, 0.5, _hip.ptr(half))

    def kick(y, a, w):
        out = torch.empty_like(y)
        _hip.call("nb_axpy_rows", _hip.ptr(y), _hip.ptr(a), _hip.ptr(half), n, w, _hip.ptr(out))
        return out

    p_half = kick(st.p, st.a_hydro, 3)
    u_half = kick(st.u, st.dudt, 1)
    if coupled._gn(cfg):
        x1, p_flow, _ = glf_device(st.x, p_half, dt_p)
    else:  # straight drift x + dt p
        x1 = _dev.empty((n, 3))
        _hip.call("nb_axpy_rows", _hip.ptr(st.x), _hip.ptr(p_half), _hip.ptr(dt_p), n, 3, _hip.ptr(x1))
        p_flow = p_half.clone()
    p_pred = kick(p_flow, st.a_hydro, 3)
    u_pred = kick(u_half, st.dudt, 1)
    order = coupled.grid_permutation(cfg, x1, st.h, st)
    counts = st.row_counts.view(torch.int64).view(n)  # two int32 counts per particle move as one
    x1, p_flow, p_pred, u_half, u_pred, a_prev, mass, h, alpha, perm, half, counts = permute_many(
        (x1, p_flow, p_pred, u_half, u_pred, st.a_hydro, st.mass, st.h, st.alpha, st.perm, half, counts), order)
    pred = replace(st, x=x1, p=p_flow, u=u_pred, mass=mass, h=h, alpha=alpha, perm=perm,
                   row_counts=counts.view(torch.int32).view(n, 2))  # fmt: skip
    new = _hydro(pred, cfg, velocity(cfg, x1, p_pred), 0.0, a_prev)

    return replace(new, p=kick(p_flow, new.a_hydro, 3), u=kick(u_half, new.dudt, 1),
                   t=st.t + float(np.max(dt_clouds)))  # fmt: skip


def cloud_state(st: GasState, k: int) -> dict:
    """Host copies of cloud k's fields in its own caller order (as GasState.to_numpy)."""
    a, b = int(st.seg_off[k]), int(st.seg_off[k + 1])
    perm = _dev.host(st.perm)[a:b] - a
    out = {}
    for name in ("x", "p", "u", "h", "rho", "a_hydro", "dudt", "alpha"):
        v = _dev.host(getattr(st, name))[a:b]
        o = np.empty_like(v)
        o[perm] = v
        out[name] = o
    return out


class Ensemble:
    """Clouds stepped together, each with its own time step and clock.

    >>> ens = Ensemble(ics, coupled.baseline_config(precision="mixed"))
    >>> for _ in range(100):
    ...     ens.step()
    >>> ens.t            # time of each cloud
    >>> ens.cloud(3)     # cloud 3's fields in its own particle order
    """

    def __init__(self, ics, cfg: coupled.CoupledConfig, cell_factor: float = 1.0):
        self.cfg = cfg
        self.state = initialise(ics, cfg, cell_factor)
        self.t = np.zeros(len(ics))
        self.sizes = np.diff(self.state.seg_off)

    @property
    def n_clouds(self) -> int:
        return len(self.sizes)

    def stable_dt(self) -> np.ndarray:
        return stable_dt_clouds(self.state, self.cfg)

    def step(self, dt: np.ndarray | None = None) -> np.ndarray:
        """One step of every cloud (by default each at its own stable step); returns the steps."""
        dt = self.stable_dt() if dt is None else np.broadcast_to(np.asarray(dt, float), (self.n_clouds,)).copy()
        self.state = step(self.state, self.cfg, dt)
        self.t += dt
        return dt

    def cloud(self, k: int) -> dict:
        return cloud_state(self.state, k)


def ensemble_ics(n_clouds: int = 64, n_min: int = 8192, n_max: int = 262144, seed: int = 20261014,
                 total_cap: int | None = None):  # fmt: skip
    """The batched-ensemble workload: n_clouds clouds of the cloud physics, particle counts
    log-uniform in [n_min, n_max], pericentres from 1000 to 4000, radii from 100 to 200, each
    cloud launched from r = 20000 on its parabolic orbit (nebula.ics.jittered_sphere_on_parabola)."""
    from ..ics import jittered_sphere_on_parabola

    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_clouds)).astype(int)
    if total_cap is not None and counts.sum() > total_cap:
        counts = np.maximum((counts * (total_cap / counts.sum())).astype(int), 64)
    r_p = np.linspace(1000.0, 4000.0, n_clouds)
    radius = np.linspace(100.0, 200.0, n_clouds)
    return [jittered_sphere_on_parabola(int(counts[k]), radius=float(radius[k]), r_p=float(r_p[k]), seed=seed + k)
            for k in range(n_clouds)]  # fmt: skip
