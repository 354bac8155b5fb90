"""Closed-form physics used by the step, oracle side (test infrastructure only).

NumPy restatements of the reference's formulas (files cited per function), plus the two
laws the workloads need that the reference does not have (power-law cooling, Balsara
switch), defined here from their published forms.
"""

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
