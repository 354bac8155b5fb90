"""Closed-form physics used by the step, oracle side (test infrastructure only).

NumPy restatements of the reference's formulas (files cited per function), plus the two
laws the workloads need that the reference does not have (power-law cooling, Balsara
switch), defined here from their published forms.
"""

import math

import numpy as np

# SI constants, CODATA 2018 (units.py:17-39 of the reference)
G = 6.67430e-11
C = 299_792_458.0
M_SUN = 1.98847e30
K_B = 1.380649e-23
M_P = 1.67262192369e-27
MU = 0.6
M_BH_MSUN = 4.3e6
T_FLOOR = 1.0e4
# free-free emissivity of fully ionized hydrogen, n_e = n_i = rho/m_p, Gaunt factor 1
# (Rybicki & Lightman 1979, Eq. 5.15b: 1.4e-27 erg cm^3 s^-1 K^-1/2 = 1.4e-40 W m^3 K^-1/2)
LAMBDA_FF = 1.4e-40

LENGTH_M = G * M_BH_MSUN * M_SUN / C**2
TIME_S = G * M_BH_MSUN * M_SUN / C**3
DENSITY_KG_M3 = M_BH_MSUN * M_SUN / LENGTH_M**3  # code mass unit = the hole's mass


def gas_u_from_temperature(t):
    """units.py:131-133: u = 3 k T / (2 mu m_p) / c^2."""
    return 1.5 * K_B * t / (MU * M_P) / C**2


def ideal_u_floor(t_floor, gamma):
    """hydro.py:155-163 (ideal branch) with eos.py:38-42: k T / ((gamma-1) mu m_p) / c^2."""
    return K_B * t_floor / ((gamma - 1.0) * MU * M_P) / C**2


def eos_ideal(rho, u, gamma):
    """eos.py:21-28: P = (gamma-1) rho u, c_s = sqrt(gamma (gamma-1) u)."""
    return (gamma - 1.0) * np.asarray(rho) * np.asarray(u), np.sqrt(gamma * (gamma - 1.0) * np.asarray(u))


# ----------------------------------------------------------------- gas + radiation EOS
SIGMA_SB = 5.670374419e-8  # W m^-2 K^-4 (units.py:26)
A_RAD = 4.0 * SIGMA_SB / C  # radiation constant, J m^-3 K^-4 (units.py:27)
T_MIN_SI, T_MAX_SI = 1.0, 1.0e13  # solver bracket in K (eos.py:82)
EOS_TOL = 1e-12  # eos.py:83


def gas_density_kg_m3(cloud_mass_msun=0.1, cloud_mass_code=1.0, bh_mass_msun=M_BH_MSUN):
    """units.py:89-110 (GasUnits.density_kg_m3): one code density unit, (cloud_mass_msun M_sun /
    cloud_mass_code) / (G M_bh / c^2)^3, in kg/m^3."""
    length = G * bh_mass_msun * M_SUN / C**2
    return cloud_mass_msun * M_SUN / cloud_mass_code / length**3


def gasrad_u_si(rho_si, t_k):
    """eos.py:86-92: u = 3 k T / (2 mu m_p) + a T^4 / rho  [J/kg]."""
    t = np.asarray(t_k, float)
    return 1.5 * K_B * t / (MU * M_P) + A_RAD * t**4 / np.asarray(rho_si)


def gasrad_pressure_si(rho_si, t_k):
    """eos.py:95-100: P = rho k T / (mu m_p) + a T^4 / 3  [Pa]."""
    t = np.asarray(t_k, float)
    return np.asarray(rho_si) * K_B * t / (MU * M_P) + A_RAD * t**4 / 3.0


def gasrad_beta_si(rho_si, t_k):
    """eos.py:103-109: beta_gas = P_gas / P."""
    t = np.asarray(t_k, float)
    p_gas = np.asarray(rho_si) * K_B * t / (MU * M_P)
    return p_gas / (p_gas + A_RAD * t**4 / 3.0)


def gamma1(beta, gamma=5.0 / 3.0):
    """eos.py:112-117: Gamma_1 = beta + (4 - 3 beta)^2 (gamma - 1) / (beta + 12 (gamma - 1)(1 - beta))."""
    b = np.asarray(beta, float)
    g = gamma - 1.0
    return b + (4.0 - 3.0 * b) ** 2 * g / (b + 12.0 * g * (1.0 - b))


def gasrad_temperature_si(rho_si, u_si, tol=EOS_TOL, max_iter=200):
    """fastphys.py:45-93 (the per-element form of eos.py:120-148): T [K] from rho [kg/m^3] and
    u [J/kg] by Newton on log T, kept inside a bracket [log 1, log 1e13] narrowed at every step
    (a Newton step leaving it, or not finite, bisects); starts from the smaller of the gas-only
    and radiation-only temperatures; stops once a step in log T is below tol. ValueError when
    u lies outside u(1 K) .. u(1e13 K), RuntimeError when an element does not converge."""
    rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho_si, float), np.shape(u_si)), float).ravel()
    u = np.ascontiguousarray(np.asarray(u_si, float)).ravel()
    shape = np.shape(u_si)
    if np.any(u < gasrad_u_si(rho, T_MIN_SI)) or np.any(u > gasrad_u_si(rho, T_MAX_SI)):
        raise ValueError("u outside the EOS bracket [1, 1e13] K")
    a1 = 1.5 * K_B / (MU * M_P)
    lt_min, lt_max = np.log(T_MIN_SI), np.log(T_MAX_SI)
    out = np.empty_like(u)
    for i in range(len(u)):
        r, uu = float(rho[i]), float(u[i])
        lo, hi = lt_min, lt_max
        t0 = min(uu / a1, (uu * r / A_RAD) ** 0.25)
        lt = math.log(min(max(t0, T_MIN_SI), T_MAX_SI))
        ok = False
        for _ in range(max_iter):
            t = math.exp(lt)
            t4 = t**4
            f = a1 * t + A_RAD * t4 / r - uu
            dfdl = a1 * t + 4.0 * A_RAD * t4 / r
            if f < 0:
                lo = lt
            if f > 0:
                hi = lt
            new = lt - f / dfdl
            if new <= lo or new >= hi or not math.isfinite(new):
                new = 0.5 * (lo + hi)
            step = abs(new - lt)
            lt = new
            if step < tol:
                ok = True
                break
        if not ok:
            raise RuntimeError("gas + radiation EOS: temperature did not converge")
        out[i] = math.exp(lt)
    return out.reshape(shape)


def gasrad_state(rho_code, u_code, density_kg_m3):
    """fastphys.py:96-111: (P [code], c_s [c], T [K], beta_gas) from code rho and u, one code
    density unit being density_kg_m3; c_s = sqrt(Gamma_1 P / rho)."""
    rho_si = np.asarray(rho_code, float) * density_kg_m3
    u_si = np.asarray(u_code, float) * C**2
    t = gasrad_temperature_si(rho_si, u_si)
    p_si = gasrad_pressure_si(rho_si, t)
    beta = gasrad_beta_si(rho_si, t)
    p_code = p_si / (density_kg_m3 * C**2)
    cs = np.sqrt(gamma1(beta) * p_code / np.asarray(rho_code, float))
    return p_code, cs, t, beta


def gasrad_u_floor(rho_code, t_floor, density_kg_m3):
    """hydro.py:155-163 (gas + radiation branch): the code u of gas at t_floor at each density."""
    return gasrad_u_si(np.asarray(rho_code, float) * density_kg_m3, t_floor) / C**2


# ----------------------------------------------------------------- cooling
def beta_cool(u, u_floor, dt, t_cool):
    """cooling.py:33-40: u_floor + (u - u_floor) exp(-dt / t_cool)."""
    with np.errstate(divide="ignore"):
        decay = np.exp(-np.asarray(dt, float) / np.asarray(t_cool, float))
    return u_floor + (u - u_floor) * decay


def cooling_time(r, beta):
    """cooling.py:28-30: beta / Omega_K = beta r^1.5."""
    return beta / np.asarray(r, float) ** -1.5


def powerlaw_coefficient(gamma):
    """k in du/dt = -k rho sqrt(u) (code units) for Lambda = LAMBDA_FF (rho/m_p)^2 sqrt(T)
    per unit volume, T = (gamma-1) mu m_p u c^2 / k_B."""
    return (TIME_S * DENSITY_KG_M3 * LAMBDA_FF * np.sqrt((gamma - 1.0) * MU * M_P * C**2 / K_B)
            / (M_P**2 * C**2))  # fmt: skip


def powerlaw_cool(u, rho, dt, u_floor, k):
    """Exact solution of du/dt = -k rho sqrt(u) at fixed rho, stopped at u_floor:
    sqrt(u) falls linearly, sqrt(u(dt)) = sqrt(u) - k rho dt / 2. Gas at or below the floor
    is left alone."""
    u = np.asarray(u, float)
    s = np.sqrt(u) - 0.5 * k * np.asarray(rho, float) * dt
    sf = np.sqrt(u_floor)
    return np.where(u > u_floor, np.where(s > sf, s * s, u_floor), u)


# ----------------------------------------------------------------- viscosity switches
def balsara_xi(div_v, curl_v):
    """viscosity.py:82-87"""
    d2 = np.asarray(div_v, float) ** 2
    c2 = np.asarray(curl_v, float) ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(d2 + c2 > 0, d2 / (d2 + c2), 0.0)


def alpha_local(h, ddivv_dt, xi, c_s, alpha_max=1.0):
    """viscosity.py:90-93"""
    a = xi * np.maximum(-np.asarray(ddivv_dt, float), 0.0)
    return np.minimum(10.0 * h * h * a / (c_s * c_s), alpha_max)


def decay_time(h, c_s, sigma=0.1):
    """viscosity.py:96-98"""
    return h / (sigma * c_s)


def alpha_update(alpha, alpha_loc, dt, tau):
    """viscosity.py:101-106"""
    decayed = (alpha + alpha_loc * dt / tau) / (1.0 + dt / tau)
    return np.where(alpha_loc > alpha, alpha_loc, decayed)


BALSARA_EPS = 1e-4


def balsara_factor(div_v, curl_v, c_s, h):
    """Balsara (1995) shear limiter f = |div v| / (|div v| + |curl v| + eps c_s / h)."""
    a = np.abs(div_v)
    return a / (a + np.abs(curl_v) + BALSARA_EPS * c_s / h)


# ----------------------------------------------------------------- Navier-Stokes / disc
def ss_viscosity(alpha, c_s, r):
    """navier_stokes.py:29-31"""
    return np.asarray(alpha) * np.asarray(c_s) ** 2 * np.asarray(r, float) ** 1.5


def viscous_stress(g, rho, nu, zeta_v=0.0):
    """navier_stokes.py:34-40"""
    rho, nu, zeta_v = (np.asarray(a, float)[..., None, None] for a in (rho, nu, zeta_v))
    div = np.trace(g, axis1=-2, axis2=-1)[..., None, None]
    return rho * ((zeta_v - 2.0 / 3.0 * nu) * div * np.eye(3) + nu * (g + np.swapaxes(g, -1, -2)))


def viscous_heating(g, nu, zeta_v=0.0):
    """navier_stokes.py:43-50"""
    t = g + np.swapaxes(g, -1, -2)
    div = np.trace(g, axis1=-2, axis2=-1)
    return (np.asarray(zeta_v) - 2.0 / 3.0 * np.asarray(nu)) * div**2 + 0.5 * np.asarray(
        nu
    ) * np.sum(t * t, axis=(-2, -1))


def viscous_dt(h, nu, c_visc=0.25):
    """navier_stokes.py:53-57"""
    nu = np.asarray(nu, float)
    with np.errstate(divide="ignore"):
        return np.where(nu > 0, c_visc * np.asarray(h, float) ** 2 / nu, np.inf)


def nu_av(alpha_av, c_s, h):
    """disc_viscosity.py:46-48"""
    return np.asarray(alpha_av) * np.asarray(c_s) * np.asarray(h) / 10.0


# ----------------------------------------------------------------- GN Hamiltonian
def _radius(x):
    return np.sqrt(np.sum(x * x, axis=-1, keepdims=True))


def gn_p_from_v(x, v):
    """gravity.py:66-77"""
    r = _radius(x)
    f = r / (r - 2.0)
    g = 2.0 * r / (r - 2.0) ** 2
    rdot = np.sum(x * v, axis=-1, keepdims=True) / r
    return f * v + g * rdot * x / r


def gn_hamiltonian(x, p):
    """gravity.py:92-102"""
    r = _radius(x)
    a, c = (r - 2.0) / r, -2.0 * (r - 2.0) / r**4
    dot = np.sum(p * x, axis=-1, keepdims=True)
    return (0.5 * a * np.sum(p * p, axis=-1, keepdims=True) + 0.5 * c * dot**2 - 1.0 / r)[..., 0]


def gn_angmom(x, p):
    """gravity.py:121-124"""
    return np.cross(x, p)
