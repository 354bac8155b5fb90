"""Block (individual power-of-two) time steps, oracle side (test infrastructure only).

Restates blocks.py of the reference in NumPy on the oracle's C passes, in the caller's particle
order: ``initialise`` (blocks.py:455-499), ``advance`` (:502-563), ``_hydro_rows`` (:242-315)
with ``_density_touched`` (:217-239), ``wanted_dt`` (:318-335), the level choice with limiter,
landing and wake-up (:338-431), the kicks and heat budget (:434-452), ``run`` (:722-739),
``ledger`` (:742-758), ``diagnostics`` (:761-774) and ``snapshot`` (:777-787). Not restated:
removal inside r_acc / beyond r_escape (:566-627, 673-719) and gravity sub-steps (:630-670),
which the product rejects as well; the pericentre flags of the removal bookkeeping are kept.
"""

import math
from dataclasses import dataclass, field

import numpy as np

from . import coupled as oc
from . import physics as ph
from .neighbours import Neighbours, tighten
from .sph import compute_density, forces_numba, gradients_numba, ns_forces_numba, ns_gradients_numba
from .sph import disc_forces_numba, velocity_gradients_numba
from .timesteps import K_MAX, allowed_level, level_for_dt, limit_rows, ticks


@dataclass
class BlockState:
    """blocks.py:104-169 (without the removal ledger and sub-step fields)."""

    n: int
    t0: float
    tick: float
    level: np.ndarray
    n_beg: np.ndarray
    n_end: np.ndarray
    x_beg: np.ndarray
    p_half: np.ndarray
    u_half: np.ndarray
    x: np.ndarray
    p: np.ndarray
    u: np.ndarray
    mass: np.ndarray
    h: np.ndarray
    alpha: np.ndarray
    a_hydro: np.ndarray
    dudt: np.ndarray
    rho: np.ndarray
    omega: np.ndarray
    vsig_max: np.ndarray
    level_offset: np.ndarray
    n_stop: int = 0
    neighbour_cache: object = None
    active_counts: list = field(default_factory=list)
    particle_steps: int = 0
    woken: int = 0
    last_pairs: object = None
    r0: np.ndarray = None
    vr_neg: np.ndarray = None
    passed_peri: np.ndarray = None
    e_radiated: float = 0.0
    luminosity: np.ndarray = None
    visc_heating: np.ndarray = None
    cond_heating: np.ndarray = None
    disk_heating: np.ndarray = None
    nu: np.ndarray = None
    heat_visc: np.ndarray = None
    heat_cond: np.ndarray = None
    heat_disk: np.ndarray = None
    heat_pdv: np.ndarray = None
    heat_cool: np.ndarray = None

    @property
    def t(self):
        return self.t0 + self.n * self.tick

    def synced(self):
        return bool(np.all(self.n_beg == self.n))


def _check(cfg):
    if cfg.r_acc is not None or cfg.r_escape is not None or cfg.respa:
        raise NotImplementedError("removal and gravity sub-steps are not restated")


def _with_neighbours(nb, rows, n):
    """blocks.py:202-214"""
    mark = np.zeros(n, bool)
    rows = np.asarray(rows, np.int64)
    mark[rows] = True
    for i in rows:
        mark[nb.index[nb.offset[i] : nb.offset[i + 1]]] = True
    return np.flatnonzero(mark)


def _density_touched(st, hc, rows, box):
    """blocks.py:217-239"""
    n = len(st.x)
    cache = st.neighbour_cache
    todo = np.arange(n) if cache is None else _with_neighbours(cache.neighbours, rows, n)
    done = np.zeros(n, bool)
    for _ in range(8):
        d = compute_density(st.x, st.mass, st.h, box=box, kernel_name=hc.kernel, eta=hc.eta, h_max=hc.h_max,
                            skin=hc.skin, rows=todo, cache=st.neighbour_cache)  # fmt: skip
        st.neighbour_cache = d.cache
        st.h[todo], st.rho[todo], st.omega[todo] = d.h, d.rho, d.omega
        done[todo] = True
        nb = tighten(d.neighbours, st.x, st.h, rows=rows)
        need = _with_neighbours(nb, rows, n)
        todo = need[~done[need]]
        if len(todo) == 0:
            return nb
    raise RuntimeError("neighbour set of the active particles did not settle")


def _hydro_rows(st, cfg, rows, v, u, dt_alpha, a_prev):
    """blocks.py:242-315"""
    n = len(st.x)
    if not cfg.hydro:
        return Neighbours(np.zeros(n + 1, np.int64), np.zeros(0, np.int64), None)
    hc = cfg.coupled.hydro
    box = None if hc.box is None else np.asarray(hc.box, float)
    if np.any(u <= 0):
        raise FloatingPointError(f"non-positive internal energy at {int(np.sum(u <= 0))} particles")
    nb = _density_touched(st, hc, rows, box)
    h, rho, omega = st.h, st.rho, st.omega
    touched = _with_neighbours(nb, rows, n)
    p_gas, cs = np.full(n, np.nan), np.full(n, np.nan)
    p_gas[touched], cs[touched] = oc.eos_state(hc, rho[touched], u[touched], st.x[touched])
    alpha = st.alpha.copy()
    if hc.viscosity and hc.switch and np.any(dt_alpha > 0):
        g = gradients_numba(st.x, v, a_prev, st.mass, h, nb, hc.kernel, rows=rows)
        a_loc = ph.alpha_local(h[rows], g.ddivv_dt, ph.balsara_xi(g.div_v, g.curl_v), cs[rows])
        alpha[rows] = ph.alpha_update(alpha[rows], a_loc, dt_alpha, ph.decay_time(h[rows], cs[rows]))
    elif hc.viscosity and not hc.switch:
        alpha[rows] = hc.alpha_const
    recon = hc.reconstruction and hc.viscosity
    grad_v = None
    if recon:
        grad_v = velocity_gradients_numba(st.x, v, st.mass, h, st.neighbour_cache.neighbours, hc.kernel,
                                          rows=touched)  # fmt: skip
    f = forces_numba(
        st.x, v, st.mass, h, rho, omega, p_gas, nb, hc.kernel,
        u=u if hc.conductivity else None, c_s=cs if hc.viscosity else None,
        alpha=alpha if hc.viscosity else None, beta=hc.beta, alpha_u=hc.alpha_u,
        cond_signal=hc.conductivity_signal, reconstruction=recon, closeness=hc.recon_closeness,
        eta_crit=1.0 / hc.eta, recon_gate=hc.recon_gate, rows=rows, grad_v=grad_v,
    )  # fmt: skip
    acc, dudt = f.acc, f.dudt
    nu = oc.physical_viscosity(hc, st.x[touched], cs[touched])
    if nu is not None:
        g = ns_gradients_numba(st.x, v, st.mass, h, rho, omega, st.neighbour_cache.neighbours, hc.kernel,
                               rows=touched)  # fmt: skip
        st.nu[touched] = nu
        stress = np.zeros_like(g)
        stress[touched] = ph.viscous_stress(g[touched], rho[touched], nu)
        acc = acc + ns_forces_numba(st.x, st.mass, h, rho, omega, stress, nb, hc.kernel, rows=rows)
        st.disk_heating[rows] = ph.viscous_heating(g[rows], st.nu[rows])
        dudt = dudt + st.disk_heating[rows]
    if hc.disc_alpha_av > 0:
        a_disc, u_disc = disc_forces_numba(st.x, v, st.mass, h, rho, omega, cs, hc.disc_alpha_av, nb, hc.kernel,
                                           rows=rows)  # fmt: skip
        acc = acc + a_disc
        st.disk_heating[rows] = u_disc
        dudt = dudt + u_disc
        st.nu[rows] = ph.nu_av(hc.disc_alpha_av, cs[rows], h[rows])
    st.alpha = alpha
    st.a_hydro[rows] = acc
    st.dudt[rows] = dudt
    st.vsig_max[rows] = f.vsig_max if hc.viscosity else cs[rows]
    st.visc_heating[rows] = f.dudt_visc
    st.cond_heating[rows] = f.dudt_cond
    if hc.isothermal_cs is not None:
        for arr in (st.dudt, st.visc_heating, st.cond_heating, st.disk_heating):
            arr[rows] = 0.0
    return nb


def wanted_dt(st, cfg, rows):
    """blocks.py:318-335"""
    cc = cfg.coupled
    if not cfg.hydro:
        return cc.eta_orb * np.linalg.norm(st.x[rows], axis=1) ** 1.5
    hc = cc.hydro
    h = st.h[rows]
    dt = hc.courant * h / st.vsig_max[rows]
    amag = np.linalg.norm(st.a_hydro[rows], axis=1)
    with np.errstate(divide="ignore"):
        dt = np.minimum(dt, np.where(amag > 0, hc.c_force * np.sqrt(h / amag), np.inf))
    dt = np.minimum(dt, ph.viscous_dt(h, st.nu[rows]))
    if cc.hole:
        dt = np.minimum(dt, cc.eta_orb * np.linalg.norm(st.x[rows], axis=1) ** 1.5)
    return dt


def _land(st, k):
    """blocks.py:383-393"""
    if st.n >= st.n_stop:
        return k
    while True:
        over = st.n + ticks(np.minimum(k, K_MAX)) > st.n_stop
        if not np.any(over):
            return k
        k = np.where(over, k + 1, k)


def _add_heat(st, rows, dt):
    """blocks.py:434-440"""
    visc, cond, disk = st.visc_heating[rows], st.cond_heating[rows], st.disk_heating[rows]
    st.heat_visc[rows] += dt * visc
    st.heat_cond[rows] += dt * cond
    st.heat_disk[rows] += dt * disk
    st.heat_pdv[rows] += dt * (st.dudt[rows] - visc - cond - disk)


def _cut_to(st, j, end_new, level):
    """blocks.py:420-431"""
    st.level[j] = level
    d_len = (end_new - st.n_end[j]) * st.tick
    st.p_half[j] += 0.5 * d_len[:, None] * st.a_hydro[j]
    st.u_half[j] += 0.5 * d_len * st.dudt[j]
    _add_heat(st, j, 0.5 * d_len)
    st.n_end[j] = end_new
    st.woken += len(j)


def _wake(st, cfg, rows, nb):
    """blocks.py:396-417 (_wake, _wake_targets, _cut_steps)"""
    need = np.full(len(st.level), -1, np.int64)
    for i in rows:
        for j in nb.index[nb.offset[i] : nb.offset[i + 1]]:
            need[j] = max(need[j], st.level[i] - cfg.max_gap)
    inactive = st.n_end > st.n
    j = np.flatnonzero(inactive & (need > st.level))
    if len(j) == 0:
        return
    t_new = ticks(need[j])
    end_new = (st.n // t_new + 1) * t_new
    st.level[j] = need[j]
    cut = end_new < st.n_end[j]
    _cut_to(st, j[cut], end_new[cut], need[j[cut]])


def _choose_levels(st, cfg, rows, nb):
    """blocks.py:367-380"""
    k = level_for_dt(wanted_dt(st, cfg, rows), cfg.dt_max) + st.level_offset[rows]
    k = allowed_level(st.n, st.level[rows], k)
    level = st.level.copy()
    level[rows] = k
    if cfg.limiter:
        limit_rows(level, rows, nb.offset, nb.index, cfg.max_gap)
    level[rows] = _land(st, level[rows])
    if np.any(level[rows] > K_MAX):
        raise ValueError(f"a particle needs a level beyond K_MAX = {K_MAX} (investigate it)")
    st.level = level
    if cfg.limiter:
        _wake(st, cfg, rows, nb)


def _open(st, rows):
    """blocks.py:443-452"""
    length = ticks(st.level[rows]) * st.tick
    st.n_beg[rows] = st.n
    st.n_end[rows] = st.n + ticks(st.level[rows])
    st.x_beg[rows] = st.x[rows]
    st.p_half[rows] = st.p[rows] + 0.5 * length[:, None] * st.a_hydro[rows]
    st.u_half[rows] = st.u[rows] + 0.5 * length * st.dudt[rows]
    _add_heat(st, rows, 0.5 * length)


def _flags(st, cfg, rows):
    """blocks.py:680-689: pericentre flags of the active particles (no removal radii)."""
    cc = cfg.coupled
    x, p = st.x[rows], st.p[rows]
    r = np.linalg.norm(x, axis=1)
    vr = np.sum(x * oc.velocity(cc, x, p), axis=1)
    turned = st.vr_neg[rows] & (vr > 0) & (r < 0.5 * st.r0[rows])
    st.passed_peri[rows] |= turned
    st.vr_neg[rows] = vr < 0


def initialise(x, v, u, mass, h_guess, cfg, t_stop, t0=0.0, level_offset=None):
    """blocks.py:455-499"""
    _check(cfg)
    cc = cfg.coupled
    hc = cc.hydro
    x = np.array(x, float)
    n = len(x)
    p = oc.p_from_v(cc, x, np.asarray(v, float))
    tick = cfg.dt_max / 2.0**K_MAX
    z3 = np.zeros((n, 3))
    st = BlockState(
        n=0, t0=t0, tick=tick, level=np.zeros(n, np.int64), n_beg=np.zeros(n, np.int64),
        n_end=np.zeros(n, np.int64), x_beg=x.copy(), p_half=p.copy(), u_half=np.array(u, float) * np.ones(n),
        x=x, p=p, u=np.array(u, float) * np.ones(n), mass=np.array(mass, float) * np.ones(n),
        h=np.array(h_guess, float) * np.ones(n), alpha=np.full(n, hc.alpha_init if hc.switch else hc.alpha_const),
        a_hydro=z3.copy(), dudt=np.zeros(n), rho=np.zeros(n), omega=np.ones(n), vsig_max=np.zeros(n),
        level_offset=np.zeros(n, np.int64) if level_offset is None else np.asarray(level_offset, np.int64),
        n_stop=int(round((t_stop - t0) / tick)),
    )  # fmt: skip
    st.r0 = np.linalg.norm(x, axis=1)
    st.vr_neg = np.sum(x * oc.velocity(cc, x, p), axis=1) < 0
    st.passed_peri = np.zeros(n, bool)
    st.luminosity = np.zeros(n)
    st.visc_heating, st.cond_heating, st.disk_heating, st.nu = (np.zeros(n) for _ in range(4))
    st.heat_visc, st.heat_cond, st.heat_disk, st.heat_pdv, st.heat_cool = (np.zeros(n) for _ in range(5))
    rows = np.arange(n)
    nb = _hydro_rows(st, cfg, rows, oc.velocity(cc, x, p), st.u, np.zeros(n), None)
    st.level[:] = 0
    _choose_levels(st, cfg, rows, nb)
    _open(st, rows)
    return st


def advance(st, cfg):
    """blocks.py:502-563 (no removal, no sub-steps)"""
    _check(cfg)
    cc = cfg.coupled
    n_next = int(np.min(st.n_end))
    s = (n_next - st.n_beg) * st.tick
    half_len = 0.5 * (st.n_end - st.n_beg) * st.tick
    x, p_flow = oc.gravity_flow(cc, st.x_beg, st.p_half, s)
    rows = np.flatnonzero(st.n_end == n_next)
    lag = (s - half_len)[:, None]
    p_pred = p_flow + lag * st.a_hydro
    if math.isfinite(cc.cooling_beta):
        r_mid = 0.5 * (np.linalg.norm(st.x_beg, axis=1) + np.linalg.norm(x, axis=1))
        u_half_c = ph.beta_cool(st.u_half, oc.floor_energy(cc.hydro, st.rho, cc.t_floor), s,
                                ph.cooling_time(r_mid, cc.cooling_beta))  # fmt: skip
        dm = st.mass[rows] * (st.u_half[rows] - u_half_c[rows])
        st.e_radiated += float(np.sum(dm))
        st.luminosity[rows] = dm / (2.0 * half_len[rows])
        u_pred = u_half_c + (s - half_len) * st.dudt
        st.heat_cool[rows] += st.u_half[rows] - u_half_c[rows]
        st.u_half[rows] = u_half_c[rows]
    else:
        u_pred = st.u_half + (s - half_len) * st.dudt
    a_prev = st.a_hydro.copy()
    st.n = n_next
    st.x = x
    st.u = u_pred
    v_eval = oc.velocity(cc, x, p_pred)
    nb = _hydro_rows(st, cfg, rows, v_eval, u_pred, 2.0 * half_len[rows], a_prev)
    p_pred[rows] = p_flow[rows] + half_len[rows, None] * st.a_hydro[rows]
    st.u[rows] = st.u_half[rows] + half_len[rows] * st.dudt[rows]
    _add_heat(st, rows, half_len[rows])
    st.p = p_pred
    if cc.hole:
        _flags(st, cfg, rows)
    _choose_levels(st, cfg, rows, nb)
    _open(st, rows)
    st.last_pairs = nb
    st.active_counts.append(len(rows))
    st.particle_steps += len(rows)
    return rows


def run(st, cfg, t_end, callback=None):
    """blocks.py:722-739"""
    n_end = int(round((t_end - st.t0) / st.tick))
    st.n_stop = n_end
    fresh = np.flatnonzero((st.n_beg == st.n) & (st.n_end > n_end))
    if len(fresh):
        st.level[fresh] = _land(st, st.level[fresh])
        if np.any(st.level[fresh] > K_MAX):
            raise ValueError(f"t_end needs a level beyond K_MAX = {K_MAX}")
        _open(st, fresh)
    if np.any(st.n_end > n_end):
        raise ValueError("some particles' current steps already run past t_end")
    while st.n < n_end:
        advance(st, cfg)
        if callback is not None:
            callback(st)
    return st


def ledger(st, cfg):
    """blocks.py:742-758 (nothing removed)"""
    cc = cfg.coupled
    h = ph.gn_hamiltonian(st.x, st.p) if oc._gn(cc) else 0.5 * np.sum(st.p * st.p, axis=1)
    am = ph.gn_angmom(st.x, st.p)
    return {
        "energy": float(np.sum(st.mass * (h + st.u))) + st.e_radiated,
        "e_radiated": st.e_radiated,
        "abs_h": float(np.sum(st.mass * np.abs(h))),
        "angmom": np.sum(st.mass[:, None] * am, axis=0),
        "momentum": np.sum(st.mass[:, None] * st.p, axis=0),
        "mass": float(np.sum(st.mass)),
    }


def snapshot(st, cfg, t):
    """blocks.py:777-787"""
    n_t = (t - st.t0) / st.tick
    if np.any(n_t < st.n_beg - 1e-6) or np.any(n_t > st.n_end + 1e-6):
        raise ValueError("snapshot time outside some particle's current step")
    s = np.maximum((t - st.t0) - st.n_beg * st.tick, 0.0)
    x, p_flow = oc.gravity_flow(cfg.coupled, st.x_beg, st.p_half, s)
    lag = s - 0.5 * (st.n_end - st.n_beg) * st.tick
    return x, p_flow + lag[:, None] * st.a_hydro
