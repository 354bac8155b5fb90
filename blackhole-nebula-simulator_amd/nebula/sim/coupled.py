"""Gas in the hole's gravity on the GPU: SPH coupled to the GN Hamiltonian through canonical
momenta (coupled.py of the reference).

The call surface is the reference's (coupled.py:62-270): ``CoupledConfig``, ``GasState``,
``initialise``, ``step``, ``stable_dt``, ``ledger``. The state lives in GPU memory for the whole
run (torch tensors; ``GasState.to_numpy()`` copies it out), and every operation of the step is
a HIP kernel: half kicks, GN generalized-leapfrog flow, cooling, neighbour lists, density,
equation of state, tightening, viscosity switch, forces, time-step minima and the ledger.
The host only sequences the launches and reads back a few scalars (list sizes, rebuild
counts, the step's minima).

Particle order. Before each density pass the particles are put in the order of the neighbour
grid (cells in raster order; nb_grid_order), so that every gathered neighbour row comes out
sorted and the lanes of a wavefront work on neighbouring particles. ``GasState.perm`` maps the
stored order back to the caller's (``to_numpy`` returns the caller's order). Sums over a row run
in the stored order, so results equal those of the reference's arithmetic applied to the
particles in that order; the oracle applies the same ordering. The mixed-precision mode orders
by the segmented grid of an ensemble (nb_grid_order_seg), the cloud being its one segment: the
same lists, with z cells a quarter of the x-y cell for a tighter candidate walk.

Each step (coupled.py:238-258): hydro half-kick; GN flow over dt; cooling in the middle of
the split; hydro forces at the new positions with the predicted velocity; second half-kick.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import torch

from .. import _dev, _hip
from ..physics import units
from .fastphys import gn_v_from_p_device
from .hydro import HydroConfig, floor_energy, gasrad_state_device
from .integrator import glf_device
from .neighbours import IndexBuffer, grid_order, grid_order_seg, permute, permute_many
from .sph import (
    COND_VELOCITY,
    hydro_density_device,
    hydro_density_mixed_device,
    density_device,
    disc_forces_device,
    forces_device,
    gradients_device,
    ns_forces_device,
    ns_gradients_device,
    velocity_gradients_device,
)

H_MAX_HOLE = 200.0
BALSARA_EPS = 1e-4


@dataclass(frozen=True)
class CoupledConfig:
    hydro: HydroConfig = HydroConfig(conductivity_signal=COND_VELOCITY, alpha_u=0.05, h_max=H_MAX_HOLE)
    hole: bool = True
    eta_orb: float = 0.00125
    wrong_coupling: bool = False
    gravity: str = "gn"
    cooling_beta: float = math.inf
    t_floor: float = units.T_FLOOR
    # addition of this implementation: optically thin power-law cooling (rho^2 T^1/2)
    cooling_powerlaw: bool = False


def baseline_config(**hydro_overrides) -> CoupledConfig:
    """The plain method of the cloud workloads: cubic spline, eta 1.2, Monaghan viscosity with
    alpha = 1 and beta = 2 (no switch, no reconstruction), no conductivity, h_max 200."""
    hc = HydroConfig(kernel="cubic", eta=1.2, h_max=H_MAX_HOLE, switch=False, alpha_const=1.0, beta=2.0,
                     reconstruction=False, conductivity=False)  # fmt: skip
    if hydro_overrides:
        hc = replace(hc, **hydro_overrides)
    return CoupledConfig(hydro=hc)


@dataclass
class GasState:
    t: float
    x: torch.Tensor
    p: torch.Tensor
    u: torch.Tensor
    mass: torch.Tensor
    h: torch.Tensor
    alpha: torch.Tensor
    a_hydro: torch.Tensor = field(default=None)
    dudt: torch.Tensor = field(default=None)
    rho: torch.Tensor = field(default=None)
    c_s: torch.Tensor = field(default=None)
    vsig_max: torch.Tensor = field(default=None)
    neighbour_cache: object = field(default=None)
    e_radiated_dev: torch.Tensor = field(default=None)  # (1,) device accumulator
    visc_heating: torch.Tensor = field(default=None)
    nu: torch.Tensor = field(default=None)
    omega: torch.Tensor = field(default=None)
    iterations: torch.Tensor = field(default=None)
    flags: torch.Tensor = field(default=None)
    m_uniform: float = 0.0
    list_capacity: int = 0
    builds: int = 0
    n_search_entries: int = 0  # search-list entries of the last density pass (incl. self)
    n_tight_entries: int = 0  # tightened-list entries of the last force pass (incl. self)
    perm: torch.Tensor = field(default=None)  # caller index of each stored particle (int64)
    row_cap: int = 0
    # learned-h inputs kept from the last hydro pass, and its statistics
    div_v: torch.Tensor = field(default=None)
    curl_v: torch.Tensor = field(default=None)
    nneigh: torch.Tensor = field(default=None)  # tight-list length per particle (float64)
    dt_prev: float = 0.0
    learned_fallbacks_dev: torch.Tensor = field(default=None)  # (1,) int64: particles that fell back
    # ensembles of independent clouds (nebula.sim.ensemble): system of each particle, their ranges
    cloud: torch.Tensor = field(default=None)
    seg_off: np.ndarray = field(default=None)
    cell_factor: float = 1.0  # neighbour-grid cell edge relative to the default (mixed path; autotuned)
    # storage of the tight lists, reused from step to step (neighbours.IndexBuffer)
    list_buffer: object = field(default=None, compare=False, repr=False)
    # mixed mode: (N, 2) int32 search-row and extras counts of the last density pass (sizes the next)
    row_counts: torch.Tensor = field(default=None, compare=False, repr=False)

    @property
    def learned_fallbacks(self) -> int:
        return 0 if self.learned_fallbacks_dev is None else int(self.learned_fallbacks_dev.item())

    @property
    def e_radiated(self) -> float:
        return 0.0 if self.e_radiated_dev is None else float(self.e_radiated_dev.item())

    # pickling (checkpoints, nebula.sim.runio): device tensors travel as host arrays, bit for bit,
    # and come back on the current GPU; the tight-list buffer is scratch and is not kept
    def __getstate__(self):
        state = {}
        for k, v in self.__dict__.items():
            if k == "list_buffer":
                v = None
            elif isinstance(v, torch.Tensor):
                v = ("__device_tensor__", v.detach().cpu().numpy())
            state[k] = v
        return state

    def __setstate__(self, state):
        for k, v in state.items():
            if isinstance(v, tuple) and len(v) == 2 and v[0] == "__device_tensor__":
                v = torch.from_numpy(np.ascontiguousarray(v[1])).to(_hip.torch_device())
            self.__dict__[k] = v

    def to_numpy(self) -> dict:
        """Host copies in the caller's particle order."""
        out = {}
        perm = _dev.host(self.perm)
        for k in ("x", "p", "u", "mass", "h", "alpha", "a_hydro", "dudt", "rho", "c_s", "vsig_max", "visc_heating",
                  "nu", "omega", "iterations", "flags"):  # fmt: skip
            v = getattr(self, k)
            if v is None:
                out[k] = None
                continue
            a = _dev.host(v)
            b = np.empty_like(a)
            b[perm] = a
            out[k] = b
        out["t"] = self.t
        out["e_radiated"] = self.e_radiated
        return out


def _gn(cfg: CoupledConfig) -> bool:
    if cfg.gravity not in ("gn",):
        raise ValueError(f"unsupported gravity {cfg.gravity!r} (the GPU step implements the GN potential)")
    if cfg.wrong_coupling:
        raise ValueError("the wrong-coupling comparison is not implemented on the GPU")
    return cfg.hole and cfg.gravity == "gn"


def _axpy(y, a, s, out=None):
    out = torch.empty_like(y) if out is None else out
    _hip.call("nb_axpy", _hip.ptr(y), _hip.ptr(a), float(s), y.numel(), _hip.ptr(out))
    return out


def _axpy_dev(y, a, s_dev, scale):
    """_axpy with s = scale * s_dev[0] read on the device (a recorded step's dt)."""
    out = torch.empty_like(y)
    _hip.call("nb_axpy_dev", _hip.ptr(y), _hip.ptr(a), _hip.ptr(s_dev), float(scale), y.numel(), _hip.ptr(out))
    return out


def _kick2(p, a, u, dudt, dts, scale):
    """(p + s a, u + s dudt) with s = scale * dts[0] read on the device, in one launch."""
    n = int(u.numel())
    if p.numel() != 3 * n or a.numel() != 3 * n or dudt.numel() != n:
        raise ValueError("a kick pairs (N, 3) momenta with N energies")
    p_out = torch.empty_like(p)
    u_out = torch.empty_like(u)
    _hip.call("nb_kick2_dev", _hip.ptr(p), _hip.ptr(a), _hip.ptr(u), _hip.ptr(dudt), _hip.ptr(dts), float(scale),
              n, _hip.ptr(p_out), _hip.ptr(u_out))  # fmt: skip
    return p_out, u_out


def _step_dt(fixed):
    """The device (dt, dt_prev) of a step recorded for replay (fixed.dts), or None."""
    return None if fixed is None else getattr(fixed, "dts", None)


def _full(n, v):
    t = _dev.empty(n)
    _hip.call("nb_fill", _hip.ptr(t), float(v), n)
    return t


def _search_lengths(cfg: CoupledConfig, h):
    n = int(h.shape[0])
    out = _dev.empty(n)
    _hip.call("nb_search_lengths", _hip.ptr(h), float(cfg.hydro.h_max), float(cfg.hydro.skin), float(np.inf), n,
              _hip.ptr(out))  # fmt: skip
    return out


_ONE_CLOUD: dict = {}


def one_cloud(n: int, device):
    """(cloud, seg_off) of a single system as one segment: the mixed-precision passes always run on
    the segmented grid (an ensemble's, here with one segment), the FP64 passes on the plain grid
    that the oracle restates."""
    key = (int(n), str(device))
    if key not in _ONE_CLOUD:
        _ONE_CLOUD.clear()
        _ONE_CLOUD[key] = (torch.zeros(n, dtype=torch.int32, device=device), np.array([0, n], np.int64))
    return _ONE_CLOUD[key]


def _segments(cfg: CoupledConfig, st: GasState | None, n: int, device):
    """The (cloud, seg_off) the grid passes of this state use, or (None, None) for the plain grid."""
    if st is not None and st.cloud is not None:
        return st.cloud, st.seg_off
    if cfg.hydro.precision == "mixed":
        return one_cloud(n, device)
    return None, None


def grid_permutation(cfg: CoupledConfig, x, h, st: GasState | None = None, fixed=None):
    """Order of the neighbour grid the next density pass will build (search lengths from h).
    fixed: the fixed-capacity passes (no host synchronisation; see _hydro)."""
    if fixed is not None:
        return fixed.order(x, _search_lengths(cfg, h), 1.0 if st is None else st.cell_factor)
    cloud, seg_off = _segments(cfg, st, int(x.shape[0]), x.device)
    if cloud is not None:
        return grid_order_seg(x, _search_lengths(cfg, h), cloud, seg_off, 1.0 if st is None else st.cell_factor)
    return grid_order(x, _search_lengths(cfg, h), None, 1.0 if st is None else st.cell_factor)


def velocity(cfg: CoupledConfig, x, p):
    """v = dH/dp (coupled.py:104-109)."""
    if not _gn(cfg):
        return p
    return gn_v_from_p_device(x, p)


def p_from_v(cfg: CoupledConfig, x: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Canonical momentum from velocity (gravity.py:66-77), on the host at set-up."""
    if not _gn(cfg):
        return np.array(v, float)
    r = np.sqrt(np.sum(x * x, axis=-1, keepdims=True))
    f = r / (r - 2.0)
    g = 2.0 * r / (r - 2.0) ** 2
    rdot = np.sum(x * v, axis=-1, keepdims=True) / r
    return f * v + g * rdot * x / r


def gravity_flow(cfg: CoupledConfig, x, p, dt, fixed=None):
    """coupled.py:131-142 (GN generalized leapfrog, or a straight drift without the hole). fixed:
    failures are counted in fixed.status[6] instead of raising (no host synchronisation)."""
    dts = _step_dt(fixed)
    if _gn(cfg) and fixed is not None:
        n = int(x.shape[0])
        x1, p1 = _dev.empty((n, 3)), _dev.empty((n, 3))
        iters = _dev.empty(n, torch.int32)
        dt_p = None
        if dts is not None:  # the recorded step's dt, one value per particle
            dt_p = _dev.empty(n)
            _hip.call("nb_expand_seg", _hip.ptr(dts), _hip.ptr(fixed.cloud), n, 1.0, _hip.ptr(dt_p))
        _hip.call("nb_glf_step", _hip.ptr(x), _hip.ptr(p), _hip.ptr(dt_p), float(dt), n, _hip.ptr(x1), _hip.ptr(p1),
                  _hip.ptr(iters), _hip.ptr(fixed.status[6:]))  # fmt: skip
    elif _gn(cfg):
        x1, p1, _ = glf_device(x, p, float(dt))
    else:
        x1 = _axpy(x, p, dt) if dts is None else _axpy_dev(x, p, dts, 1.0)
        p1 = p.clone()
    if cfg.hydro.box is not None:
        raise ValueError("periodic boxes are not supported by the coupled GPU step")
    return x1, p1


def _hydro(st: GasState, cfg: CoupledConfig, v, dt_alpha: float, a_prev, h_start=None, fixed=None) -> GasState:
    """coupled.py:145-195 on device tensors. h_start: the Newton solve's starting h (default:
    the previous h, as in the reference); the search lists always come from the previous h.
    fixed: the fixed-capacity passes (nebula.sim.captured: a step recorded for replay, or one
    checked after the fact by nebula.sim.stepper): one search round, no host synchronisation;
    what does not fit goes to fixed.status."""
    hc = cfg.hydro
    if hc.isothermal_cs is not None or hc.eos not in ("ideal", "gasrad"):
        raise ValueError("the GPU step implements the ideal-gas and gas + radiation equations of state")
    n = int(st.x.shape[0])
    buf = st.list_buffer if st.list_buffer is not None else IndexBuffer()
    counts = None
    if fixed is not None:
        counts = _dev.empty((n, 2), torch.int32)
        d, n_search = fixed.density(st, cfg, h_start, counts)
    elif hc.precision == "mixed":
        counts = _dev.empty((n, 2), torch.int32)
        cloud, seg_off = _segments(cfg, st, n, st.x.device)
        d, n_search = hydro_density_mixed_device(st.x, st.mass, st.h.clone(), hc.kernel, hc.eta, 1e-4, hc.h_max,
                                                 hc.skin, st.m_uniform, h_start_d=h_start, cloud_d=cloud,
                                                 seg_off=seg_off, cell_factor=st.cell_factor,
                                                 index_buffer=buf, row_hint=st.row_counts,
                                                 row_counts=counts)  # fmt: skip
    else:
        if st.cloud is not None or st.cell_factor != 1.0:
            raise ValueError("ensembles and tuned grids run in the mixed-precision mode")
        counts = _dev.empty((n, 2), torch.int32)
        d, n_search = hydro_density_device(st.x, st.mass, st.h.clone(), hc.kernel, hc.eta, 1e-4, hc.h_max, hc.skin,
                                           st.m_uniform, hc.precision, h_start_d=h_start, index_buffer=buf,
                                           row_hint=st.row_counts, row_counts=counts)  # fmt: skip
    p_gas = _dev.empty(n)
    cs = _dev.empty(n)
    nbad = torch.zeros(1, dtype=torch.int32, device=st.x.device) if fixed is None else fixed.status[7:]
    if hc.eos == "gasrad":
        # u <= 0 lies below the solver bracket: counted with the bracket failures
        _hip.call("nb_eos_ideal", _hip.ptr(d.rho), _hip.ptr(st.u), float(hc.gamma), n, _hip.ptr(p_gas),
                  _hip.ptr(cs), _hip.ptr(nbad))  # fmt: skip
        bad = int(nbad.item()) if fixed is None else 0
        if bad:
            raise FloatingPointError(f"non-positive internal energy at {bad} particles")
        p_gas, cs, _ = gasrad_state_device(hc, d.rho, st.u, None if fixed is None else fixed.status)
    else:
        _hip.call("nb_eos_ideal", _hip.ptr(d.rho), _hip.ptr(st.u), float(hc.gamma), n, _hip.ptr(p_gas),
                  _hip.ptr(cs), _hip.ptr(nbad))  # fmt: skip
        bad = int(nbad.item()) if fixed is None else 0  # a recorded step reports it in its status
        if bad:
            raise FloatingPointError(f"non-positive internal energy at {bad} particles")
    nb = d.neighbours  # tightened at the final h
    alpha = st.alpha
    g = None
    if hc.viscosity and hc.switch and dt_alpha > 0:
        g = gradients_device(st.x, v, a_prev, st.mass, d.h, nb, hc.kernel, precision=hc.precision)
        alpha = _dev.empty(n)
        _hip.call("nb_alpha_switch", _hip.ptr(g.div_v), _hip.ptr(g.curl_v), _hip.ptr(g.ddivv_dt), _hip.ptr(cs),
                  _hip.ptr(d.h), _hip.ptr(st.alpha), float(dt_alpha), 1.0, 0.1, n, _hip.ptr(alpha))  # fmt: skip
    elif hc.viscosity and not hc.switch:
        if hc.balsara:
            g = gradients_device(st.x, v, None, st.mass, d.h, nb, hc.kernel, precision=hc.precision)
            alpha = _dev.empty(n)
            _hip.call("nb_alpha_balsara", _hip.ptr(g.div_v), _hip.ptr(g.curl_v), _hip.ptr(cs), _hip.ptr(d.h),
                      float(hc.alpha_const), BALSARA_EPS, n, _hip.ptr(alpha))  # fmt: skip
        else:
            alpha = _full(n, hc.alpha_const)
    if g is None and _learned(cfg) is not None:
        # div v and |curl v| for the learned start's features
        g = gradients_device(st.x, v, None, st.mass, d.h, nb, hc.kernel, precision=hc.precision)
    shock = hc.shock()
    gv = None
    if shock.reconstruction:
        gv = torch.zeros((n, 3, 3), dtype=torch.float64, device=st.x.device)
        velocity_gradients_device(st.x, v, st.mass, d.h, nb, hc.kernel, gv)
    f = forces_device(st.x, v, st.mass, d.h, d.rho, d.omega, p_gas, nb, hc.kernel,
                      st.u if hc.conductivity else None, cs if hc.viscosity else None,
                      alpha if hc.viscosity else None, shock, gv_d=gv, precision=hc.precision)  # fmt: skip
    acc, dudt = f.acc, f.dudt
    nu = None
    if (hc.ns_nu > 0) + (hc.ns_alpha > 0) + (hc.disc_alpha_av > 0) > 1:
        raise ValueError("set only one of ns_nu, ns_alpha, disc_alpha_av")
    if hc.ns_nu > 0 or hc.ns_alpha > 0:
        gns = torch.zeros((n, 3, 3), dtype=torch.float64, device=st.x.device)
        ns_gradients_device(st.x, v, st.mass, d.h, d.rho, d.omega, nb, hc.kernel, gns)
        nu = _dev.empty(n)
        stress = _dev.empty((n, 3, 3))
        heat = _dev.empty(n)
        _hip.call("nb_ns_stress", _hip.ptr(gns), _hip.ptr(d.rho), _hip.ptr(cs), _hip.ptr(st.x), float(hc.ns_alpha),
                  float(hc.ns_nu), n, _hip.ptr(nu), _hip.ptr(stress), _hip.ptr(heat))  # fmt: skip
        a_ns = ns_forces_device(st.x, st.mass, d.h, d.rho, d.omega, stress, nb, hc.kernel)
        _hip.call("nb_add", _hip.ptr(acc), _hip.ptr(a_ns), 3 * n, _hip.ptr(acc))
        _hip.call("nb_add", _hip.ptr(dudt), _hip.ptr(heat), n, _hip.ptr(dudt))
    if hc.disc_alpha_av > 0:
        a_disc, u_disc = disc_forces_device(st.x, v, st.mass, d.h, d.rho, d.omega, cs, hc.disc_alpha_av, nb,
                                            hc.kernel)  # fmt: skip
        _hip.call("nb_add", _hip.ptr(acc), _hip.ptr(a_disc), 3 * n, _hip.ptr(acc))
        _hip.call("nb_add", _hip.ptr(dudt), _hip.ptr(u_disc), n, _hip.ptr(dudt))
        nu = _dev.empty(n)
        _hip.call("nb_nu_av", _hip.ptr(cs), _hip.ptr(d.h), float(hc.disc_alpha_av), n, _hip.ptr(nu))
    return replace(
        st, h=d.h, alpha=alpha, a_hydro=acc, dudt=dudt, rho=d.rho, c_s=cs,
        vsig_max=f.vsig_max if hc.viscosity else cs, visc_heating=f.dudt_visc, nu=nu, omega=d.omega,
        iterations=d.iterations, flags=d.flags, builds=d.builds,
        n_search_entries=n_search,
        n_tight_entries=nb.n_total,
        div_v=None if g is None else g.div_v, curl_v=None if g is None else g.curl_v,
        nneigh=nb.length.double(), list_buffer=buf, row_counts=counts,
    )  # fmt: skip


def _learned(cfg: CoupledConfig):
    """The learned-h network of the configuration, or None."""
    hc = cfg.hydro
    return hc.learned_model if (hc.learned_h or getattr(hc.learned_model, "collect", None)) else None


def initialise(x, v, u, mass, h_guess, cfg: CoupledConfig, t: float = 0.0, uniform_mass: bool = True) -> GasState:
    """Start from positions and velocities, p computed from v (coupled.py:198-210). Equal masses
    are passed to the pair kernels as one number unless ``uniform_mass`` is False (the per-particle
    form an ensemble of clouds uses)."""
    x = np.array(x, float)
    n = len(x)
    v = np.array(v, float)
    p = p_from_v(cfg, x, v)
    hc = cfg.hydro
    mass_h = np.array(mass, float) * np.ones(n)
    m_uniform = float(mass_h[0]) if uniform_mass and n and np.all(mass_h == mass_h[0]) and mass_h[0] > 0 else 0.0
    st = GasState(
        t, _dev.f64(x), _dev.f64(p), _dev.f64(np.array(u, float) * np.ones(n)), _dev.f64(mass_h),
        _dev.f64(np.array(h_guess, float) * np.ones(n)),
        _dev.f64(np.full(n, hc.alpha_init if hc.switch else hc.alpha_const)),
        e_radiated_dev=torch.zeros(1, dtype=torch.float64, device=_hip.torch_device()), m_uniform=m_uniform,
    )  # fmt: skip
    order = grid_permutation(cfg, st.x, st.h)
    st = replace(st, x=permute(st.x, order), p=permute(st.p, order), u=permute(st.u, order),
                 mass=permute(st.mass, order), h=permute(st.h, order), alpha=permute(st.alpha, order),
                 perm=order.to(torch.int64))  # fmt: skip
    return _hydro(st, cfg, velocity(cfg, st.x, st.p), 0.0, None)


def stable_dt(st: GasState, cfg: CoupledConfig) -> float:
    """Shared step: min over particles of the Courant, force, viscous and orbit criteria
    (coupled.py:213-226). The minima are taken on the device (exact in any order); r^1.5 of the
    smallest radius equals the smallest r^1.5 (pow is monotone)."""
    n = int(st.x.shape[0])
    out = np.zeros(4)
    _hip.call("nb_dt_minima", _hip.ptr(st.h), _hip.ptr(st.vsig_max), _hip.ptr(st.a_hydro), _hip.ptr(st.nu), 0.25,
              _hip.ptr(st.x), n, out.ctypes.data_as(_hip._PD))  # fmt: skip
    return dt_from_minima(out, cfg, st.nu is not None)


def dt_from_minima(out, cfg: CoupledConfig, viscous: bool) -> float:
    """stable_dt from the four device minima (min h / vsig, min sqrt(h / |a|), min c h^2 / nu,
    min |x|)."""
    hc = cfg.hydro
    dt = hc.courant * float(out[0])
    if np.isfinite(out[1]):
        dt = min(dt, hc.c_force * float(out[1]))
    if viscous:
        dt = min(dt, float(out[2]))
    if cfg.hole:
        dt = min(dt, cfg.eta_orb * float(out[3]) ** 1.5)  # scalar (libm) pow, as the oracle
    return float(dt)


def step(st: GasState, cfg: CoupledConfig, dt: float, fixed=None) -> GasState:
    """One Strang-split step (coupled.py:238-258), with optional power-law cooling in the same
    place of the split as the reference's beta cooling. fixed: the fixed-capacity passes of
    nebula.sim.stepper (no host synchronisation; the learned start's pairs are handed to
    fixed.pairs for the caller to log once the step is known to be good)."""
    hc = cfg.hydro
    n = int(st.x.shape[0])
    half = 0.5 * dt
    dts = _step_dt(fixed)
    if dts is not None and (math.isfinite(cfg.cooling_beta) or (hc.viscosity and hc.switch)):
        raise ValueError("a recorded step covers power-law cooling and the constant or Balsara viscosity")
    model = _learned(cfg)
    gate = getattr(model, "gate", None) if hc.learned_h else None
    decided = getattr(fixed, "use_net", None)  # a pipelined driver consults the gate itself
    if gate is not None and dts is not None and decided is None:
        raise ValueError("a recorded step runs the learned start on every replay: set the model's gate to None")
    # decided before any launch of this step (from the counts of the steps before the last)
    use_net = hc.learned_h and (bool(decided) if decided is not None else (gate is None or gate.decide()))

    def kick(y, a):
        return _axpy(y, a, half) if dts is None else _axpy_dev(y, a, dts, 0.5)

    def kick2(p, a, u, dudt):  # a momentum and an energy kick of the same half step
        if dts is None:
            return kick(p, a), kick(u, dudt)
        return _kick2(p, a, u, dudt, dts, 0.5)

    p_half, u_half = kick2(st.p, st.a_hydro, st.u, st.dudt)
    x1, p_flow = gravity_flow(cfg, st.x, p_half, dt, fixed)
    e_rad = st.e_radiated_dev.clone()
    if math.isfinite(cfg.cooling_beta) or cfg.cooling_powerlaw:
        u_floor = units.ideal_u_floor(cfg.t_floor, hc.gamma)
        e_part = _dev.empty(n)
        if cfg.cooling_powerlaw and hc.eos != "ideal":
            raise ValueError("the power-law cooling law is defined for the ideal-gas equation of state")
        if math.isfinite(cfg.cooling_beta):
            u_cool = _dev.empty(n)
            if hc.eos == "gasrad":  # the floor at each particle's density (hydro.py:155-163)
                _hip.call("nb_cool_beta_floors", _hip.ptr(u_half), _hip.ptr(st.x), _hip.ptr(x1), _hip.ptr(st.mass),
                          _hip.ptr(floor_energy(hc, st.rho, cfg.t_floor)), float(dt), float(cfg.cooling_beta), n,
                          _hip.ptr(u_cool), _hip.ptr(e_part))  # fmt: skip
            else:
                _hip.call("nb_cool_beta", _hip.ptr(u_half), _hip.ptr(st.x), _hip.ptr(x1), _hip.ptr(st.mass),
                          float(u_floor), float(dt), float(cfg.cooling_beta), n, _hip.ptr(u_cool),
                          _hip.ptr(e_part))  # fmt: skip
            _hip.call("nb_sum_into", _hip.ptr(e_part), n, _hip.ptr(e_rad))
            u_half = u_cool
        if cfg.cooling_powerlaw:
            u_cool = _dev.empty(n)
            k_cool = float(units.powerlaw_cooling_coefficient(hc.gamma))
            if dts is None:
                _hip.call("nb_cool_powerlaw", _hip.ptr(u_half), _hip.ptr(st.rho), _hip.ptr(st.mass), float(u_floor),
                          k_cool, float(dt), n, _hip.ptr(u_cool), _hip.ptr(e_part))  # fmt: skip
            else:
                _hip.call("nb_cool_powerlaw_dev", _hip.ptr(u_half), _hip.ptr(st.rho), _hip.ptr(st.mass),
                          float(u_floor), k_cool, _hip.ptr(dts), n, _hip.ptr(u_cool), _hip.ptr(e_part))  # fmt: skip
            _hip.call("nb_sum_into", _hip.ptr(e_part), n, _hip.ptr(e_rad))
            u_half = u_cool
    order = grid_permutation(cfg, x1, st.h, st, fixed)
    if dts is None:
        p_pred, u_pred = kick2(p_flow, st.a_hydro, u_half, st.dudt)
        kicks = None
    else:  # the predictor kick of a device dt is formed by the permutation, of the moved values
        p_pred = u_pred = None
        kicks = {2: (p_flow, st.a_hydro), 4: (u_half, st.dudt)}
    fields = [x1, p_flow, p_pred, u_half, u_pred, st.a_hydro, st.mass, st.h, st.alpha, st.perm]
    counts = st.row_counts
    if counts is not None:
        fields.append(counts.view(torch.int64).view(n))  # two int32 counts per particle move as one
    extra = []
    if model is not None and (use_net or model.collect is not None):
        # the learned start's inputs from the last hydro pass, in the new order
        extra = [t for t in (st.rho, st.div_v, st.curl_v, st.nneigh)]
    moved = permute_many(fields + [t for t in extra if t is not None], order, kicks=kicks, s_dev=dts, scale=0.5)
    x1, p_flow, p_pred, u_half, u_pred, a_prev, mass, h, alpha, perm = moved[:10]
    if counts is not None:
        counts = moved[10].view(torch.int32).view(n, 2)
    pred = replace(st, t=st.t + dt, x=x1, p=p_flow, u=u_pred, mass=mass, h=h, alpha=alpha, perm=perm,
                   row_counts=counts)  # fmt: skip
    h_start = feat = None
    fallbacks = None
    if model is not None and (use_net or model.collect is not None):
        it = iter(moved[len(fields):])
        rho_p, div_p, curl_p, nn_p = (None if t is None else next(it) for t in extra)
        feat = model.features(h, rho_p, div_p, curl_p, nn_p, x1, u_pred, dt, st.dt_prev, stride=1, dts=dts)
        if use_net:
            h_start = model.h_start(h, feat)
            fallbacks = model.last_fallbacks
    new = _hydro(pred, cfg, velocity(cfg, x1, p_pred), dt, a_prev, h_start=h_start, fixed=fixed)
    if gate is not None and not getattr(fixed, "reads_back", False):  # else the driver observes it
        gate.observe((new.iterations > 1).sum(), n, use_net)
    if fixed is not None:
        fixed.pairs = (model, feat, h, new.h)
    else:
        log_pairs(cfg, model, feat, h, new.h)
    p1, u1 = kick2(p_flow, new.a_hydro, u_half, new.dudt)
    return replace(new, p=p1, u=u1, e_radiated_dev=e_rad,
                   dt_prev=float(dt), learned_fallbacks_dev=fallbacks)


def log_pairs(cfg: CoupledConfig, model, feat, h_prev, h_final) -> None:
    """The learned start's (features, converged h) pairs of a step: into the model's PairLog, and
    to its online training when the start is in use."""
    if model is None or feat is None:
        return
    if model.collect is not None:
        s = model.collect.stride
        model.collect.add(feat[::s].contiguous() if s > 1 else feat, model.targets(h_prev, h_final, s))
    if model.online is not None and cfg.hydro.learned_h:
        s = model.online.stride
        model.observe(feat[::s].contiguous() if s > 1 else feat, model.targets(h_prev, h_final, s))


def ledger(st: GasState, cfg: CoupledConfig) -> dict:
    """Conserved totals (coupled.py:261-270): E = sum m (H + u) + E_radiated, sum m |H|, and the
    angular momentum sum m (x cross p). Device sums, compensated over the block partials."""
    n = int(st.x.shape[0])
    out = np.zeros(6)
    _hip.call("nb_ledger", _hip.ptr(st.x), _hip.ptr(st.p), _hip.ptr(st.u), _hip.ptr(st.mass), None, int(_gn(cfg)), n,
              out.ctypes.data_as(_hip._PD))  # fmt: skip
    e_rad = st.e_radiated
    return {"energy": float(out[0]) + e_rad, "e_radiated": e_rad, "abs_h": float(out[1]), "angmom": out[2:5].copy()}
