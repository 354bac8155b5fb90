"""Shared-step GN + SPH stepper, oracle side (test infrastructure only).

Restates coupled.py:131-270 (``gravity_flow``, ``_hydro``, ``initialise``, ``stable_dt``,
``_kick``, ``step``, ``ledger``) in NumPy on top of the C restatements. ``cfg`` is any object
with the fields of the product's ``CoupledConfig`` / ``HydroConfig`` (read by attribute).
Neighbour lists are rebuilt at every density pass (no cache), and before each density pass the
particles are put in the order of the GPU implementation's neighbour grid (``grid_order``), the
two policies the GPU path follows; ``State.perm`` maps back to the caller's order.
"""

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import physics as ph
from .integrator import glf_step_many, gn_v_from_p
from .neighbours import grid_order, tighten
from .sph import (
    compute_density,
    disc_forces_numba,
    forces_numba,
    gradients_numba,
    ns_forces_numba,
    ns_gradients_numba,
    velocity_gradients_numba,
)


@dataclass
class State:
    """coupled.py:78-95 (GasState)"""

    t: float
    x: np.ndarray
    p: np.ndarray
    u: np.ndarray
    mass: np.ndarray
    h: np.ndarray
    alpha: np.ndarray
    a_hydro: np.ndarray = field(default=None)
    dudt: np.ndarray = field(default=None)
    rho: np.ndarray = field(default=None)
    c_s: np.ndarray = field(default=None)
    vsig_max: np.ndarray = field(default=None)
    e_radiated: float = 0.0
    visc_heating: np.ndarray = field(default=None)
    nu: np.ndarray = field(default=None)
    omega: np.ndarray = field(default=None)
    iterations: np.ndarray = field(default=None)
    flags: np.ndarray = field(default=None)
    builds: int = 0
    perm: np.ndarray = field(default=None)

    def caller_order(self, name):
        """Field ``name`` in the caller's particle order."""
        a = getattr(self, name)
        out = np.empty_like(a)
        out[self.perm] = a
        return out


def _search_lengths(hc, h):
    """sph.py:371 (open box)."""
    return np.minimum(np.minimum(h, hc.h_max) * hc.skin, np.inf)


def _gn(cfg):
    """coupled.py:98-101"""
    if cfg.gravity not in ("gn",):
        raise ValueError(f"unsupported gravity {cfg.gravity!r}")
    return cfg.hole and cfg.gravity == "gn"


def velocity(cfg, x, p):
    """coupled.py:104-109"""
    return gn_v_from_p(x, p) if _gn(cfg) else p


def p_from_v(cfg, x, v):
    """coupled.py:112-114"""
    return ph.gn_p_from_v(x, v) if _gn(cfg) else np.array(v, float)


def gravity_flow(cfg, x, p, dt):
    """coupled.py:131-142"""
    if _gn(cfg):
        x1, p1, _ = glf_step_many(x, p, dt)
    else:
        x1, p1 = x + np.reshape(dt, (-1, 1)) * p, np.array(p, float)
    if cfg.hydro.box is not None:
        x1 = np.mod(x1, np.asarray(cfg.hydro.box, float))
    return x1, p1


def eos_state(hc, rho, u, x=None):
    """hydro.py:121-139 (ideal, isothermal and radiation-limit branches)."""
    if hc.isothermal_cs is not None:
        if hc.iso_r0 is not None:
            cs = hc.isothermal_cs * (np.linalg.norm(x, axis=1) / hc.iso_r0) ** (-hc.iso_q)
            return cs**2 * np.asarray(rho), cs
        return hc.isothermal_cs**2 * np.asarray(rho), np.full_like(rho, hc.isothermal_cs)
    if np.any(u <= 0):
        raise FloatingPointError(f"non-positive internal energy at {int(np.sum(u <= 0))} particles")
    if hc.eos == "ideal":
        return ph.eos_ideal(rho, u, hc.gamma)
    if hc.eos == "gasrad":
        p, cs, _, _ = ph.gasrad_state(rho, u, _density_kg_m3(hc))
        return p, cs
    if hc.eos == "radiation":
        return np.asarray(rho) * np.asarray(u) / 3.0, np.sqrt(4.0 * np.asarray(u) / 9.0)
    raise ValueError(f"unsupported equation of state {hc.eos!r}")


def _density_kg_m3(hc):
    """units.py:89-110: the code density unit of the configuration's gas mass scale."""
    gu = hc.gas_units
    return ph.gas_density_kg_m3(gu.cloud_mass_msun, gu.cloud_mass_code, gu.bh.mass_msun)


def floor_energy(hc, rho, t_floor):
    """hydro.py:155-163: the code u of the cooling floor at each particle's density."""
    if hc.eos == "gasrad":
        return ph.gasrad_u_floor(rho, t_floor, _density_kg_m3(hc))
    if hc.eos == "ideal":
        return np.full_like(np.asarray(rho, float), ph.ideal_u_floor(t_floor, hc.gamma))
    raise ValueError(f"no temperature floor for equation of state {hc.eos!r}")


def physical_viscosity(hc, x, c_s):
    """hydro.py:221-230"""
    if (hc.ns_nu > 0) + (hc.ns_alpha > 0) + (hc.disc_alpha_av > 0) > 1:
        raise ValueError("set only one of ns_nu, ns_alpha, disc_alpha_av")
    if hc.ns_nu > 0:
        return np.full(len(x), hc.ns_nu)
    if hc.ns_alpha > 0:
        return ph.ss_viscosity(hc.ns_alpha, c_s, np.linalg.norm(x, axis=1))
    return None


def _hydro(st: State, cfg, v, dt_alpha: float, a_prev) -> State:
    """coupled.py:145-195"""
    hc = cfg.hydro
    box = None if hc.box is None else np.asarray(hc.box, float)
    d = compute_density(st.x, st.mass, st.h, box=box, kernel_name=hc.kernel, eta=hc.eta,
                        h_max=hc.h_max, skin=hc.skin)  # fmt: skip
    if np.any(st.u <= 0):
        raise FloatingPointError(f"non-positive internal energy at {int(np.sum(st.u <= 0))} particles")
    p_gas, cs = eos_state(hc, d.rho, st.u, st.x)
    nb = tighten(d.neighbours, st.x, d.h)
    alpha = st.alpha
    if hc.viscosity and hc.switch and dt_alpha > 0:
        g = gradients_numba(st.x, v, a_prev, st.mass, d.h, nb, hc.kernel)
        a_loc = ph.alpha_local(d.h, g.ddivv_dt, ph.balsara_xi(g.div_v, g.curl_v), cs)
        alpha = ph.alpha_update(alpha, a_loc, dt_alpha, ph.decay_time(d.h, cs))
    elif hc.viscosity and not hc.switch:
        alpha = np.full(len(st.x), hc.alpha_const)
        if hc.balsara:
            g = gradients_numba(st.x, v, None, st.mass, d.h, nb, hc.kernel)
            alpha = hc.alpha_const * ph.balsara_factor(g.div_v, g.curl_v, cs, d.h)
    recon = hc.reconstruction and hc.viscosity
    grad_v = velocity_gradients_numba(st.x, v, st.mass, d.h, nb, hc.kernel) if recon else None
    f = forces_numba(
        st.x, v, st.mass, d.h, d.rho, d.omega, p_gas, nb, hc.kernel,
        u=st.u if hc.conductivity else None, c_s=cs if hc.viscosity else None,
        alpha=alpha if hc.viscosity else None, beta=hc.beta, alpha_u=hc.alpha_u,
        cond_signal=hc.conductivity_signal, reconstruction=recon, closeness=hc.recon_closeness,
        eta_crit=1.0 / hc.eta, recon_gate=hc.recon_gate, grad_v=grad_v,
    )  # fmt: skip
    acc, dudt = f.acc, f.dudt
    nu = physical_viscosity(hc, st.x, cs)
    if nu is not None:
        g = ns_gradients_numba(st.x, v, st.mass, d.h, d.rho, d.omega, nb, hc.kernel)
        stress = ph.viscous_stress(g, d.rho, nu)
        acc = acc + ns_forces_numba(st.x, st.mass, d.h, d.rho, d.omega, stress, nb, hc.kernel)
        dudt = dudt + ph.viscous_heating(g, nu)
    if hc.disc_alpha_av > 0:
        a_disc, u_disc = disc_forces_numba(st.x, v, st.mass, d.h, d.rho, d.omega, cs,
                                           hc.disc_alpha_av, nb, hc.kernel)  # fmt: skip
        acc, dudt = acc + a_disc, dudt + u_disc
        nu = ph.nu_av(hc.disc_alpha_av, cs, d.h)
    visc_heating = f.dudt_visc
    if hc.isothermal_cs is not None:
        dudt, visc_heating = np.zeros(len(st.x)), np.zeros(len(st.x))
    return replace(
        st, h=d.h, alpha=alpha, a_hydro=acc, dudt=dudt, rho=d.rho, c_s=cs,
        vsig_max=f.vsig_max if hc.viscosity else cs, visc_heating=visc_heating, nu=nu,
        omega=d.omega, iterations=d.iterations, flags=d.flags, builds=d.builds,
    )  # fmt: skip


def initialise(x, v, u, mass, h_guess, cfg, t=0.0) -> State:
    """coupled.py:198-210"""
    x = np.array(x, float)
    n = len(x)
    v = np.array(v, float)
    p = p_from_v(cfg, x, v)
    hc = cfg.hydro
    u = np.array(u, float) * np.ones(n)
    mass = np.array(mass, float) * np.ones(n)
    h = np.array(h_guess, float) * np.ones(n)
    o = grid_order(x, _search_lengths(hc, h))
    st = State(t, x[o], p[o], u[o], mass[o], h[o], np.full(n, hc.alpha_init if hc.switch else hc.alpha_const),
               perm=o.astype(np.int64))  # fmt: skip
    return _hydro(st, cfg, velocity(cfg, st.x, st.p), 0.0, None)


def stable_dt(st: State, cfg) -> float:
    """coupled.py:213-226"""
    hc = cfg.hydro
    dt = hc.courant * np.min(st.h / st.vsig_max)
    amag = np.linalg.norm(st.a_hydro, axis=1)
    if np.any(amag > 0):
        dt = min(dt, hc.c_force * np.min(np.sqrt(st.h[amag > 0] / amag[amag > 0])))
    if st.nu is not None:
        dt = min(dt, float(np.min(ph.viscous_dt(st.h, st.nu))))
    if cfg.hole:
        r = np.linalg.norm(st.x, axis=1)
        dt = min(dt, cfg.eta_orb * float(np.min(r)) ** 1.5)  # = min(r**1.5): pow is monotone
    return float(dt)


def step(st: State, cfg, dt: float) -> State:
    """coupled.py:238-258, with power-law cooling as a second cooling law (same place in the
    split, same bookkeeping into e_radiated)."""
    hc = cfg.hydro
    p_half = st.p + 0.5 * dt * st.a_hydro
    u_half = st.u + 0.5 * dt * st.dudt
    x1, p_flow = gravity_flow(cfg, st.x, p_half, dt)
    e_rad = st.e_radiated
    if math.isfinite(cfg.cooling_beta):
        r_mid = 0.5 * (np.linalg.norm(st.x, axis=1) + np.linalg.norm(x1, axis=1))
        u_floor = floor_energy(hc, st.rho, cfg.t_floor)
        u_cool = ph.beta_cool(u_half, u_floor, dt, ph.cooling_time(r_mid, cfg.cooling_beta))
        e_rad += float(np.sum(st.mass * (u_half - u_cool)))
        u_half = u_cool
    if cfg.cooling_powerlaw:
        if hc.eos != "ideal":
            raise ValueError("the power-law cooling law is defined for the ideal-gas equation of state")
        u_floor = ph.ideal_u_floor(cfg.t_floor, hc.gamma)
        u_cool = ph.powerlaw_cool(u_half, st.rho, dt, u_floor, ph.powerlaw_coefficient(hc.gamma))
        e_rad += float(np.sum(st.mass * (u_half - u_cool)))
        u_half = u_cool
    p_pred = p_flow + 0.5 * dt * st.a_hydro
    u_pred = u_half + 0.5 * dt * st.dudt
    o = grid_order(x1, _search_lengths(hc, st.h))
    x1, p_flow, p_pred, u_half, u_pred, a_prev = (a[o] for a in (x1, p_flow, p_pred, u_half, u_pred, st.a_hydro))
    pred = replace(st, t=st.t + dt, x=x1, p=p_flow, u=u_pred, mass=st.mass[o], h=st.h[o], alpha=st.alpha[o],
                   perm=st.perm[o])  # fmt: skip
    new = _hydro(pred, cfg, velocity(cfg, x1, p_pred), dt, a_prev)
    return replace(new, p=p_flow + 0.5 * dt * new.a_hydro, u=u_half + 0.5 * dt * new.dudt,
                   e_radiated=e_rad)  # fmt: skip


def ledger(st: State, cfg) -> dict:
    """coupled.py:261-270"""
    h = ph.gn_hamiltonian(st.x, st.p) if _gn(cfg) else 0.5 * np.sum(st.p * st.p, axis=1)
    return {
        "energy": float(np.sum(st.mass * (h + st.u))) + st.e_radiated,
        "e_radiated": st.e_radiated,
        "abs_h": float(np.sum(st.mass * np.abs(h))),
        "angmom": np.sum(st.mass[:, None] * (ph.gn_angmom(st.x, st.p) if _gn(cfg) else
                                             np.cross(st.x, st.p)), axis=0),
    }  # fmt: skip
