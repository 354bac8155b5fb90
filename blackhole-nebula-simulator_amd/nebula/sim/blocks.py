"""Block (individual power-of-two) time steps, device-resident (blocks.py of the reference).

Every particle has its own step [n_beg, n_end] on the int64 timeline of ``timesteps`` and goes
through the shared-step kick-flow-kick of ``coupled`` (blocks.py:1-60): an opening half kick at
n_beg, the gravity flow of (x_beg, p_half) over the elapsed time, and at n_end the density and
forces of the active particles and the closing half kick. Between its sync points a particle is
only predicted, for its neighbours' sums. A system step (``advance``) goes to the earliest step
end: the particles ending there are the active rows. Their density is computed together with
all their neighbours' from the predicted positions (``_density_touched``, on neighbour lists
reused while provably complete), forces for the active rows only; new levels follow from the
step criteria, the level rules, the neighbour limiter, landing on the stop time and the wake-up
of neighbours with too long a step (Saitoh & Makino 2009; Durier & Dalla Vecchia 2012).

The state lives in device tensors in the caller's particle order; the passes are the row-subset
forms of the FP64 HIP passes (density, tightening, gradients, forces, Navier-Stokes and disc
viscosity) and the kicks, predictions and level arithmetic are HIP kernels (csrc/blocks.hip)
that follow the reference's arithmetic, so a run is the FP64 oracle's to the bit. Not
implemented (rejected at set-up): removal inside r_acc / beyond r_escape, gravity sub-steps,
periodic boxes, the isothermal equation of state, the Balsara limiter, power-law cooling and
mixed precision.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from .. import _dev, _hip
from .coupled import CoupledConfig, _gn, p_from_v, velocity
from .hydro import floor_energy, gasrad_state_device
from .integrator import GLFConvergenceError, _raise_on_failure
from .neighbours import Neighbours, tighten
from .sph import (
    compute_density,
    disc_forces_device,
    forces_device,
    gradients_device,
    ns_forces_device,
    ns_gradients_device,
    velocity_gradients_device,
)
from .timesteps import K_MAX, limit_rows_device

__all__ = ["BlockConfig", "BlockState", "initialise", "advance", "run", "ledger", "diagnostics", "snapshot",
           "GLFConvergenceError"]  # fmt: skip


@dataclass(frozen=True)
class BlockConfig:
    """blocks.py:84-101"""

    coupled: CoupledConfig = CoupledConfig()
    dt_max: float = 4096.0
    limiter: bool = True
    max_gap: int = 2
    hydro: bool = True
    r_acc: float | None = None
    r_escape: float | None = None
    respa: bool = False
    respa_eta: float = 0.05
    removal_sync: bool = True


class BlockState:
    """blocks.py:104-169 on the device: per-particle tensors (caller order), the tick counter n
    and the bookkeeping of the run."""

    _ARRAYS = ("level", "n_beg", "n_end", "x_beg", "p_half", "u_half", "x", "p", "u", "mass", "h", "alpha", "a_hydro",
               "dudt", "rho", "omega", "vsig_max", "level_offset", "r0", "vr_neg", "passed_peri", "luminosity",
               "visc_heating", "cond_heating", "disk_heating", "nu", "heat_visc", "heat_cond", "heat_disk",
               "heat_pdv", "heat_cool")  # fmt: skip

    def __init__(self, n_now: int, t0: float, tick: float, n_stop: int):
        self.n = int(n_now)
        self.t0 = float(t0)
        self.tick = float(tick)
        self.n_stop = int(n_stop)
        self.neighbour_cache = None
        self.active_counts: list = []
        self.particle_steps = 0
        self.woken = 0
        self.last_pairs = None
        self.timers = {"flow": 0.0, "hydro": 0.0, "levels": 0.0}
        self.e_rad_dev = None

    @property
    def t(self) -> float:
        return self.t0 + self.n * self.tick

    @property
    def e_radiated(self) -> float:
        return float(self.e_rad_dev.item())

    def synced(self) -> bool:
        """True when every particle is at a step boundary."""
        return bool(torch.all(self.n_beg == self.n).item())

    def to_numpy(self) -> dict:
        """Host copies of the per-particle arrays (caller order), with t and e_radiated."""
        out = {k: _dev.host(getattr(self, k)) for k in self._ARRAYS}
        out["t"] = self.t
        out["e_radiated"] = self.e_radiated
        return out


def _check(cfg: BlockConfig) -> None:
    hc = cfg.coupled.hydro
    if cfg.r_acc is not None or cfg.r_escape is not None:
        raise NotImplementedError("removal inside r_acc / beyond r_escape is not implemented on the GPU")
    if cfg.respa:
        raise NotImplementedError("gravity sub-steps are not implemented on the GPU")
    if not cfg.hydro:
        return
    if hc.box is not None:
        raise NotImplementedError("periodic boxes are not implemented by the GPU block steps")
    if hc.isothermal_cs is not None or hc.eos not in ("ideal", "gasrad"):
        raise NotImplementedError("the GPU block steps implement the ideal-gas and gas + radiation equations of state")
    if hc.balsara or hc.precision != "fp64" or hc.learned_h:
        raise NotImplementedError("the GPU block steps run the reference's FP64 physics (no Balsara, mixed, learned h)")
    if cfg.coupled.cooling_powerlaw:
        raise NotImplementedError("the block steps cool with the beta law only (blocks.py:514-529)")


def _i64(x) -> torch.Tensor:
    return _dev.i64(x)


def _flow(cc: CoupledConfig, x, p, s):
    """Gravity-only flow over per-particle times s (coupled.py:131-142)."""
    n = int(x.shape[0])
    x1, p1 = _dev.empty((n, 3)), _dev.empty((n, 3))
    if _gn(cc):
        iters = _dev.empty(n, torch.int32)
        nfail = torch.zeros(1, dtype=torch.int32, device=x.device)
        _hip.call("nb_glf_step", _hip.ptr(x), _hip.ptr(p), _hip.ptr(s), 0.0, n, _hip.ptr(x1), _hip.ptr(p1),
                  _hip.ptr(iters), _hip.ptr(nfail))  # fmt: skip
        if int(nfail.item()) > 0:
            _raise_on_failure(_dev.host(iters))
    else:
        _hip.call("nb_axpy_rows", _hip.ptr(x), _hip.ptr(p), _hip.ptr(s), n, 3, _hip.ptr(x1))
        p1.copy_(p)
    return x1, p1


def _with_neighbours(nb: Neighbours, rows: torch.Tensor, n: int) -> torch.Tensor:
    """The rows together with every particle in their lists, sorted (blocks.py:202-214)."""
    mark = torch.zeros(n, dtype=torch.uint8, device=rows.device)
    _hip.call("nb_block_mark", _hip.ptr(rows), int(rows.numel()), _hip.ptr(nb.d_offset), _hip.ptr(nb.d_index),
              _hip.ptr(mark))  # fmt: skip
    return torch.nonzero(mark).squeeze(1)


def _density_touched(st: BlockState, hc, rows: torch.Tensor) -> Neighbours:
    """Density, h and Omega of the rows and all their interacting neighbours at the current
    positions, from the reused lists first, adding neighbours until none is missing
    (blocks.py:217-239). Returns the rows' pair lists."""
    n = int(st.x.shape[0])
    cache = st.neighbour_cache
    todo = torch.arange(n, device=rows.device) if cache is None else _with_neighbours(cache.neighbours, rows, n)
    done = torch.zeros(n, dtype=torch.bool, device=rows.device)
    for _ in range(8):
        d = compute_density(st.x, st.mass, st.h, kernel_name=hc.kernel, eta=hc.eta, h_max=hc.h_max, skin=hc.skin,
                            cache=st.neighbour_cache, rows=todo)  # fmt: skip
        st.neighbour_cache = d.cache
        st.h[todo], st.rho[todo], st.omega[todo] = d.h, d.rho, d.omega
        done[todo] = True
        nb = tighten(d.neighbours, st.x, st.h, rows=rows)
        need = _with_neighbours(nb, rows, n)
        todo = need[~done[need]]
        if int(todo.numel()) == 0:
            return nb
    raise RuntimeError("neighbour set of the active particles did not settle")


def _eos_rows(hc, rho, u, idx, n):
    """P and c_s at the particles idx (NaN elsewhere, so that a stray read shows at once)."""
    p_gas = torch.full((n,), float("nan"), dtype=torch.float64, device=rho.device)
    cs = torch.full((n,), float("nan"), dtype=torch.float64, device=rho.device)
    r_t, u_t = rho[idx].contiguous(), u[idx].contiguous()
    m = int(idx.numel())
    if hc.eos == "gasrad":
        p_t, c_t, _ = gasrad_state_device(hc, r_t, u_t)
    else:
        p_t, c_t = _dev.empty(m), _dev.empty(m)
        nbad = torch.zeros(1, dtype=torch.int32, device=rho.device)
        _hip.call("nb_eos_ideal", _hip.ptr(r_t), _hip.ptr(u_t), float(hc.gamma), m, _hip.ptr(p_t), _hip.ptr(c_t),
                  _hip.ptr(nbad))  # fmt: skip
    p_gas[idx], cs[idx] = p_t, c_t
    return p_gas, cs


def _add(a, b):
    out = torch.empty_like(a)
    _hip.call("nb_add", _hip.ptr(a), _hip.ptr(b), a.numel(), _hip.ptr(out))
    return out


def _hydro_rows(st: BlockState, cfg: BlockConfig, rows, v, u, dt_alpha, a_prev) -> Neighbours:
    """Density (rows and their neighbours), switch and forces (rows only) at the current
    positions, velocities v and internal energies u of all particles (blocks.py:242-315).
    Updates st's per-particle tensors; returns the rows' interacting-pair lists."""
    n = int(st.x.shape[0])
    if not cfg.hydro:
        return Neighbours.from_device(torch.zeros(n + 1, dtype=torch.int64, device=st.x.device),
                                      torch.zeros(1, dtype=torch.int32, device=st.x.device), None)  # fmt: skip
    hc = cfg.coupled.hydro
    nonpos = int((u <= 0).sum().item())
    if nonpos:
        raise FloatingPointError(f"non-positive internal energy at {nonpos} particles")
    nb = _density_touched(st, hc, rows)
    h, rho, omega = st.h, st.rho, st.omega
    touched = _with_neighbours(nb, rows, n)
    p_gas, cs = _eos_rows(hc, rho, u, touched, n)
    nrows = int(rows.numel())
    alpha = st.alpha.clone()
    if hc.viscosity and hc.switch and bool((dt_alpha > 0).any().item()):
        g = gradients_device(st.x, v, a_prev, st.mass, h, nb, hc.kernel, rows, nrows)
        a_rows = _dev.empty(nrows)
        # the gathered operands stay referenced until the call returns (a freed temporary's
        # memory would be handed to the next one)
        cs_r, h_r, al_r, dt_r = (t.contiguous() for t in (cs[rows], h[rows], alpha[rows], dt_alpha))
        _hip.call("nb_alpha_switch_dt", _hip.ptr(g.div_v), _hip.ptr(g.curl_v), _hip.ptr(g.ddivv_dt), _hip.ptr(cs_r),
                  _hip.ptr(h_r), _hip.ptr(al_r), _hip.ptr(dt_r), 1.0, 0.1, nrows, _hip.ptr(a_rows))  # fmt: skip
        alpha[rows] = a_rows
    elif hc.viscosity and not hc.switch:
        alpha[rows] = hc.alpha_const
    shock = hc.shock()
    gv = None
    if shock.reconstruction:
        gv = torch.zeros((n, 3, 3), dtype=torch.float64, device=st.x.device)
        velocity_gradients_device(st.x, v, st.mass, h, st.neighbour_cache.neighbours, hc.kernel, gv, touched,
                                  int(touched.numel()))  # fmt: skip
    f = forces_device(st.x, v, st.mass, h, rho, omega, p_gas, nb, hc.kernel, u if hc.conductivity else None,
                      cs if hc.viscosity else None, alpha if hc.viscosity else None, shock, rows, nrows, gv)  # fmt: skip
    acc, dudt = f.acc, f.dudt
    if (hc.ns_nu > 0) + (hc.ns_alpha > 0) + (hc.disc_alpha_av > 0) > 1:
        raise ValueError("set only one of ns_nu, ns_alpha, disc_alpha_av")
    if hc.ns_nu > 0 or hc.ns_alpha > 0:
        g = torch.zeros((n, 3, 3), dtype=torch.float64, device=st.x.device)
        ns_gradients_device(st.x, v, st.mass, h, rho, omega, st.neighbour_cache.neighbours, hc.kernel, g, touched,
                            int(touched.numel()))  # fmt: skip
        stress = torch.zeros((n, 3, 3), dtype=torch.float64, device=st.x.device)
        for idx, keep_stress in ((touched, True), (rows, False)):
            m = int(idx.numel())
            nu_t, st_t, heat_t = _dev.empty(m), _dev.empty((m, 3, 3)), _dev.empty(m)
            g_t, r_t, c_t, x_t = (t.contiguous() for t in (g[idx], rho[idx], cs[idx], st.x[idx]))
            _hip.call("nb_ns_stress", _hip.ptr(g_t), _hip.ptr(r_t), _hip.ptr(c_t), _hip.ptr(x_t), float(hc.ns_alpha),
                      float(hc.ns_nu), m, _hip.ptr(nu_t), _hip.ptr(st_t), _hip.ptr(heat_t))  # fmt: skip
            if keep_stress:
                st.nu[idx] = nu_t
                stress[idx] = st_t
            else:
                st.disk_heating[idx] = heat_t
        a_ns = ns_forces_device(st.x, st.mass, h, rho, omega, stress, nb, hc.kernel, rows, nrows)
        acc = _add(acc, a_ns)
        heat_r = st.disk_heating[rows].contiguous()
        dudt = _add(dudt, heat_r)
    if hc.disc_alpha_av > 0:
        a_disc, u_disc = disc_forces_device(st.x, v, st.mass, h, rho, omega, cs, hc.disc_alpha_av, nb, hc.kernel,
                                            rows, nrows)  # fmt: skip
        acc = _add(acc, a_disc)
        st.disk_heating[rows] = u_disc
        dudt = _add(dudt, u_disc)
        nu_r = _dev.empty(nrows)
        cs_r, h_r = cs[rows].contiguous(), h[rows].contiguous()
        _hip.call("nb_nu_av", _hip.ptr(cs_r), _hip.ptr(h_r), float(hc.disc_alpha_av), nrows, _hip.ptr(nu_r))
        st.nu[rows] = nu_r
    st.alpha = alpha
    st.a_hydro[rows] = acc
    st.dudt[rows] = dudt
    st.vsig_max[rows] = f.vsig_max if hc.viscosity else cs[rows]
    st.visc_heating[rows] = f.dudt_visc
    st.cond_heating[rows] = f.dudt_cond
    return nb


def _heat(st: BlockState):
    return (_hip.ptr(st.heat_visc), _hip.ptr(st.heat_cond), _hip.ptr(st.heat_disk), _hip.ptr(st.heat_pdv))


def _land(st: BlockState, rows: torch.Tensor, level: torch.Tensor) -> None:
    """Raise level[rows] so that steps opened now end by the stop time (blocks.py:383-393)."""
    bad = torch.zeros(1, dtype=torch.int32, device=level.device)
    _hip.call("nb_block_land", _hip.ptr(rows), int(rows.numel()), st.n, st.n_stop, _hip.ptr(level), _hip.ptr(bad))
    if int(bad.item()):
        raise ValueError(f"a particle needs a level beyond K_MAX = {K_MAX} (investigate it)")


def _choose_levels(st: BlockState, cfg: BlockConfig, rows: torch.Tensor, nb: Neighbours) -> None:
    """New levels for the active rows at tick st.n; limiter and wake-up (blocks.py:367-380)."""
    cc = cfg.coupled
    hc = cc.hydro
    level = st.level.clone()
    bad = torch.zeros(2, dtype=torch.int32, device=level.device)
    eta = cc.eta_orb if (cc.hole or not cfg.hydro) else 0.0
    _hip.call("nb_block_levels", _hip.ptr(rows), int(rows.numel()), _hip.ptr(st.x), _hip.ptr(st.h),
              _hip.ptr(st.vsig_max), _hip.ptr(st.a_hydro), _hip.ptr(st.nu), float(hc.courant), float(hc.c_force),
              float(eta), int(bool(cfg.hydro)), float(cfg.dt_max), _hip.ptr(st.level_offset), st.n, _hip.ptr(level),
              _hip.ptr(bad))  # fmt: skip
    b = bad.cpu().numpy()
    if b[0]:
        raise ValueError("time steps must be positive")
    if b[1]:
        raise ValueError(f"time step needs more than K_MAX = {K_MAX} halvings of dt_max")
    if cfg.limiter:
        limit_rows_device(level, rows, nb, cfg.max_gap)
    _land(st, rows, level)
    st.level = level
    if cfg.limiter:
        n = int(st.x.shape[0])
        need = torch.full((n,), -1, dtype=torch.int64, device=level.device)
        woken = torch.zeros(1, dtype=torch.int32, device=level.device)
        _hip.call("nb_block_wake", _hip.ptr(rows), int(rows.numel()), _hip.ptr(nb.d_offset), _hip.ptr(nb.d_index),
                  int(cfg.max_gap), n, st.n, st.tick, _hip.ptr(need), _hip.ptr(st.level), _hip.ptr(st.n_end),
                  _hip.ptr(st.a_hydro), _hip.ptr(st.dudt), _hip.ptr(st.visc_heating), _hip.ptr(st.cond_heating),
                  _hip.ptr(st.disk_heating), _hip.ptr(st.p_half), _hip.ptr(st.u_half), *_heat(st),
                  _hip.ptr(woken))  # fmt: skip
        st.woken += int(woken.item())


def _open(st: BlockState, rows: torch.Tensor) -> None:
    """Opening kicks of the rows at tick st.n (blocks.py:443-452)."""
    _hip.call("nb_block_open", _hip.ptr(rows), int(rows.numel()), st.n, st.tick, _hip.ptr(st.level),
              _hip.ptr(st.n_beg), _hip.ptr(st.n_end), _hip.ptr(st.x), _hip.ptr(st.p), _hip.ptr(st.u),
              _hip.ptr(st.a_hydro), _hip.ptr(st.dudt), _hip.ptr(st.visc_heating), _hip.ptr(st.cond_heating),
              _hip.ptr(st.disk_heating), _hip.ptr(st.x_beg), _hip.ptr(st.p_half), _hip.ptr(st.u_half),
              *_heat(st))  # fmt: skip


def _flags(st: BlockState, cfg: BlockConfig, rows: torch.Tensor) -> None:
    """Pericentre flags of the active particles (blocks.py:680-689; no removal radii)."""
    x, p = st.x[rows].contiguous(), st.p[rows].contiguous()
    v = velocity(cfg.coupled, x, p)
    r = torch.linalg.norm(x, dim=1)
    vr = (x * v).sum(dim=1)
    turned = st.vr_neg[rows] & (vr > 0) & (r < 0.5 * st.r0[rows])
    st.passed_peri[rows] = st.passed_peri[rows] | turned
    st.vr_neg[rows] = vr < 0


def initialise(x, v, u, mass, h_guess, cfg: BlockConfig, t_stop: float, t0: float = 0.0,
               level_offset=None) -> BlockState:  # fmt: skip
    """All particles start active at t0; the first steps never run past t_stop
    (blocks.py:455-499)."""
    _check(cfg)
    cc = cfg.coupled
    hc = cc.hydro
    x = np.array(x, float)
    n = len(x)
    p = p_from_v(cc, x, np.asarray(v, float))
    tick = cfg.dt_max / 2.0**K_MAX
    st = BlockState(0, t0, tick, int(round((t_stop - t0) / tick)))
    ones = np.ones(n)
    st.x = _dev.f64(x)
    st.p = _dev.f64(p)
    st.x_beg = st.x.clone()
    st.p_half = st.p.clone()
    st.u = _dev.f64(np.array(u, float) * ones)
    st.u_half = st.u.clone()
    st.mass = _dev.f64(np.array(mass, float) * ones)
    st.h = _dev.f64(np.array(h_guess, float) * ones)
    st.alpha = _dev.f64(np.full(n, hc.alpha_init if hc.switch else hc.alpha_const))
    z = np.zeros(n)
    st.a_hydro = _dev.f64(np.zeros((n, 3)))
    for k in ("dudt", "rho", "vsig_max", "luminosity", "visc_heating", "cond_heating", "disk_heating", "nu",
              "heat_visc", "heat_cond", "heat_disk", "heat_pdv", "heat_cool"):  # fmt: skip
        setattr(st, k, _dev.f64(z))
    st.omega = _dev.f64(ones)
    zi = np.zeros(n, np.int64)
    st.level, st.n_beg, st.n_end = _i64(zi), _i64(zi), _i64(zi)
    st.level_offset = _i64(zi if level_offset is None else np.asarray(level_offset, np.int64))
    st.e_rad_dev = torch.zeros(1, dtype=torch.float64, device=st.x.device)
    st.r0 = torch.linalg.norm(st.x, dim=1)
    st.vr_neg = (st.x * velocity(cc, st.x, st.p)).sum(dim=1) < 0
    st.passed_peri = torch.zeros(n, dtype=torch.bool, device=st.x.device)
    rows = torch.arange(n, device=st.x.device)
    nb = _hydro_rows(st, cfg, rows, velocity(cc, st.x, st.p), st.u, torch.zeros(n, dtype=torch.float64,
                                                                                 device=st.x.device), None)  # fmt: skip
    st.level.zero_()
    _choose_levels(st, cfg, rows, nb)
    _open(st, rows)
    return st


def advance(st: BlockState, cfg: BlockConfig) -> np.ndarray:
    """One system step, to the earliest step end (blocks.py:502-563); returns the indices of the
    active set."""
    _check(cfg)
    cc = cfg.coupled
    t0 = time.perf_counter()
    n = int(st.x.shape[0])
    n_next = int(st.n_end.min().item())
    s, half = _dev.empty(n), _dev.empty(n)
    _hip.call("nb_block_elapsed", _hip.ptr(st.n_beg), _hip.ptr(st.n_end), n_next, st.tick, n, _hip.ptr(s),
              _hip.ptr(half))  # fmt: skip
    x, p_flow = _flow(cc, st.x_beg, st.p_half, s)
    t1 = time.perf_counter()
    rows = torch.nonzero(st.n_end == n_next).squeeze(1)
    p_pred, u_pred = _dev.empty((n, 3)), _dev.empty(n)
    beta = float(cc.cooling_beta)
    cool = math.isfinite(beta)
    e_part = _dev.empty(n) if cool else None
    floors = floor_energy(cc.hydro, st.rho, cc.t_floor) if cool else None
    _hip.call("nb_block_predict", _hip.ptr(st.n_beg), _hip.ptr(st.n_end), n_next, st.tick, _hip.ptr(p_flow),
              _hip.ptr(st.a_hydro), _hip.ptr(st.u_half), _hip.ptr(st.dudt), _hip.ptr(st.x_beg), _hip.ptr(x),
              _hip.ptr(st.mass), 0.0, _hip.ptr(floors), beta if cool else math.inf, n, _hip.ptr(p_pred),
              _hip.ptr(u_pred), _hip.ptr(e_part), _hip.ptr(st.luminosity), _hip.ptr(st.heat_cool))  # fmt: skip
    if cool:
        _hip.call("nb_sum_into", _hip.ptr(e_part), n, _hip.ptr(st.e_rad_dev))
    a_prev = st.a_hydro.clone()
    st.n = n_next
    st.x = x
    st.u = u_pred
    v_eval = velocity(cc, x, p_pred)
    nb = _hydro_rows(st, cfg, rows, v_eval, u_pred, 2.0 * half[rows], a_prev)
    t2 = time.perf_counter()
    _hip.call("nb_block_close", _hip.ptr(rows), int(rows.numel()), _hip.ptr(st.n_beg), _hip.ptr(st.n_end), st.tick,
              _hip.ptr(p_flow), _hip.ptr(st.a_hydro), _hip.ptr(st.u_half), _hip.ptr(st.dudt),
              _hip.ptr(st.visc_heating), _hip.ptr(st.cond_heating), _hip.ptr(st.disk_heating), _hip.ptr(p_pred),
              _hip.ptr(st.u), *_heat(st))  # fmt: skip
    st.p = p_pred
    n_active = int(rows.numel())
    if cc.hole:
        _flags(st, cfg, rows)
    _choose_levels(st, cfg, rows, nb)
    _open(st, rows)
    st.last_pairs = nb
    st.timers["flow"] += t1 - t0
    st.timers["hydro"] += t2 - t1
    st.timers["levels"] += time.perf_counter() - t2
    st.active_counts.append(n_active)
    st.particle_steps += n_active
    return _dev.host(rows)


def run(st: BlockState, cfg: BlockConfig, t_end: float, callback=None) -> BlockState:
    """Advance until every particle is exactly at t_end (blocks.py:722-739)."""
    n_end = int(round((t_end - st.t0) / st.tick))
    st.n_stop = n_end
    fresh = torch.nonzero((st.n_beg == st.n) & (st.n_end > n_end)).squeeze(1)
    if int(fresh.numel()):
        bad = torch.zeros(1, dtype=torch.int32, device=fresh.device)
        _hip.call("nb_block_land", _hip.ptr(fresh), int(fresh.numel()), st.n, st.n_stop, _hip.ptr(st.level),
                  _hip.ptr(bad))  # fmt: skip
        if int(bad.item()):
            raise ValueError(f"t_end needs a level beyond K_MAX = {K_MAX}")
        _open(st, fresh)
    if bool((st.n_end > n_end).any().item()):
        raise ValueError("some particles' current steps already run past t_end")
    while st.n < n_end:
        advance(st, cfg)
        if callback is not None:
            callback(st)
    return st


def ledger(st: BlockState, cfg: BlockConfig) -> dict:
    """Totals as in coupled.ledger (true values when st.synced(), predicted otherwise)
    (blocks.py:742-758; nothing is removed)."""
    n = int(st.x.shape[0])
    out = np.zeros(6)
    _hip.call("nb_ledger", _hip.ptr(st.x), _hip.ptr(st.p), _hip.ptr(st.u), _hip.ptr(st.mass), None,
              int(_gn(cfg.coupled)), n, out.ctypes.data_as(_hip._PD))  # fmt: skip
    e_rad = st.e_radiated
    return {
        "energy": float(out[0]) + e_rad,
        "e_radiated": e_rad,
        "abs_h": float(out[1]),
        "angmom": out[2:5].copy(),
        "momentum": _dev.host((st.mass[:, None] * st.p).sum(dim=0)),
        "mass": float(st.mass.sum().item()),
        "removed_mass": 0.0,
        "removed_count": 0,
    }


def diagnostics(st: BlockState) -> dict:
    """Level histogram, active counts, particle-steps, wake-ups and wall time per stage
    (blocks.py:761-774)."""
    counts = np.asarray(st.active_counts)
    lv = _dev.host(st.level)
    return {
        "t": st.t,
        "level_histogram": np.bincount(lv, minlength=int(lv.max()) + 1),
        "system_steps": len(counts),
        "particle_steps": st.particle_steps,
        "mean_active": float(counts.mean()) if len(counts) else 0.0,
        "woken": st.woken,
        "timers": dict(st.timers),
    }


def snapshot(st: BlockState, cfg: BlockConfig, t: float):
    """Positions and momenta of every particle at time t between sync points (blocks.py:777-787);
    the state is untouched. Host arrays."""
    n_t = (t - st.t0) / st.tick
    nb_, ne_ = _dev.host(st.n_beg), _dev.host(st.n_end)
    if np.any(n_t < nb_ - 1e-6) or np.any(n_t > ne_ + 1e-6):
        raise ValueError("snapshot time outside some particle's current step")
    s = np.maximum((t - st.t0) - nb_ * st.tick, 0.0)
    x, p_flow = _flow(cfg.coupled, st.x_beg, st.p_half, _dev.f64(s))
    lag = s - 0.5 * (ne_ - nb_) * st.tick
    p = _dev.host(p_flow) + lag[:, None] * _dev.host(st.a_hydro)
    return _dev.host(x), p
