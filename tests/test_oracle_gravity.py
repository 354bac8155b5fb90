"""Pinning the oracle's GN gravity (CPU only): the generalized leapfrog and v = dH/dp against
closed forms, finite differences and the reference's expected values (expected-values.md §2),
plus the cooling laws against their exact solutions."""

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from helpers import check

from oracle import integrator as oi
from oracle import physics as oph


def schwarzschild_el(r_p, r_a):
    """orbits.py:27-42 of the reference: E, L from the turning points."""
    ap, aa = 1.0 - 2.0 / r_p, 1.0 - 2.0 / r_a
    l2 = (ap - aa) / (aa / r_a**2 - ap / r_p**2)
    return np.sqrt(aa * (1.0 + l2 / r_a**2)), np.sqrt(l2)


def apocentre_state(r_p, r_a):
    """orbits.py:51-62 of the reference."""
    _, ell = schwarzschild_el(r_p, r_a)
    f = r_a / (r_a - 2.0)
    x0 = np.array([r_a, 0.0, 0.0])
    v0 = np.array([0.0, ell / (f * r_a), 0.0])
    return x0, v0, f * v0


def integrate(x0, p0, dt, steps):
    x, p = x0[None].copy(), p0[None].copy()
    xs, ps = [x[0]], [p[0]]
    for _ in range(steps):
        x, p, _ = oi.glf_step_many(x, p, dt)
        xs.append(x[0])
        ps.append(p[0])
    return np.array(xs), np.array(ps)


def test_expected_orbit_constants():
    e, ell = schwarzschild_el(25.0, 5000.0)
    check("§2 L of the r_p 25 / r_a 5000 orbit", 7.35374434336, ell, rtol=1e-10)
    check("§2 E", 0.999801061330200, e, rtol=1e-11)
    x0, v0, p0 = apocentre_state(25.0, 5000.0)
    check("§2 GN energy H = (E^2-1)/2", -1.98918881503e-4, oph.gn_hamiltonian(x0, p0), rtol=1e-9)
    check("§2 apocentre speed", 1.4701605691e-3, v0[1], rtol=1e-8)


def test_v_from_p_is_dH_dp(rng):
    x = rng.normal(0, 1, (200, 3)) * 50 + np.array([100.0, 0, 0])
    p = rng.normal(0, 0.1, (200, 3))
    v = oi.gn_v_from_p(x, p)
    e = 1e-7
    fd = np.zeros_like(p)
    for c in range(3):
        dp = np.zeros(3)
        dp[c] = e
        fd[:, c] = (oph.gn_hamiltonian(x, p + dp) - oph.gn_hamiltonian(x, p - dp)) / (2 * e)
    check("v = dH/dp (central differences)", fd, v, rtol=1e-7, per_vector=True)
    check("p -> v -> p round trip", p, oph.gn_p_from_v(x, v), rtol=1e-12, per_vector=True)


def test_circular_orbit():
    """integrator test 1 (expected-values §2): r = 10, dt = 0.02 r^1.5, 200 orbits."""
    r = 10.0
    vt = (r - 2) / (r * np.sqrt(r - 3))
    x0 = np.array([r, 0.0, 0.0])
    p0 = oph.gn_p_from_v(x0[None], np.array([[0.0, vt, 0.0]]))[0]
    dt = 0.02 * r**1.5
    steps = int(np.ceil(200 * 2 * np.pi * r / vt / dt))
    xs, ps = integrate(x0, p0, dt, steps)
    h = oph.gn_hamiltonian(xs, ps)
    ell = np.linalg.norm(np.cross(xs, ps), axis=1)
    check("circular: max |dH/H| (§2: 5.6e-9)", 5.6e-9, np.max(np.abs(h - h[0])) / abs(h[0]), rtol=0.2)
    check("circular: max |dL/L| < 1e-12", 0.0, np.max(np.abs(ell - ell[0])) / ell[0], atol=1e-12)


@pytest.mark.parametrize("dt,expected", [(2.0, 1.5698e-3), (1.0, 3.9235e-4), (0.5, 9.8087e-5)])
def test_second_order_energy_error(dt, expected):
    """integrator test 2 (expected-values §2): r_p 10, r_a 100, max |dH/H| over t = 4,800."""
    x0, _, p0 = apocentre_state(10.0, 100.0)
    xs, ps = integrate(x0, p0, dt, int(round(4800.0 / dt)))
    h = oph.gn_hamiltonian(xs, ps)
    check(f"eccentric dt={dt}: max |dH/H| vs §2", expected, np.max(np.abs(h - h[0])) / abs(h[0]), rtol=0.05)


def test_reversibility():
    x0, _, p0 = apocentre_state(10.0, 100.0)
    xs, ps = integrate(x0, p0, 0.5, 2400)
    xb, pb = integrate(xs[-1], -ps[-1], 0.5, 2400)
    check("reversibility: position", x0, xb[-1], rtol=1e-8, per_vector=True)
    check("reversibility: momentum", p0, -pb[-1], rtol=1e-8, per_vector=True)


def test_failures_raise():
    x = np.array([[6.0, 0.0, 0.0]])
    p = oph.gn_p_from_v(x, np.array([[0.0, 0.4, 0.0]]))
    with pytest.raises(oi.GLFConvergenceError, match="did not converge"):
        oi.glf_step_many(x, p, 50.0)
    with pytest.raises(oi.GLFConvergenceError, match="inside r"):
        oi.glf_step_many(np.array([[2.4, 0.0, 0.0]]), p, 0.1)


# ----------------------------------------------------------------- cooling
def test_powerlaw_cooling_exact_solution():
    """du/dt = -k rho sqrt(u) above the floor: the update equals a tight numerical integration."""
    k = oph.powerlaw_coefficient(5.0 / 3.0)
    u_floor = oph.ideal_u_floor(1e4, 5.0 / 3.0)
    rho, u0 = 3e-15, 400.0 * u_floor

    def rhs(t, y):
        return [-k * rho * np.sqrt(max(y[0], 0.0)) if y[0] > u_floor else 0.0]

    t_end = 0.8 * 2.0 * (np.sqrt(u0) - np.sqrt(u_floor)) / (k * rho)  # before reaching the floor
    ode = solve_ivp(rhs, (0, t_end), [u0], rtol=1e-12, atol=1e-30).y[0, -1]
    check("power-law cooling = ODE solution", ode, oph.powerlaw_cool(u0, rho, t_end, u_floor, k), rtol=1e-9)
    check("power-law cooling stops at the floor", u_floor, oph.powerlaw_cool(u0, rho, 10 * t_end, u_floor, k))
    check("gas below the floor untouched", 0.5 * u_floor, oph.powerlaw_cool(0.5 * u_floor, rho, t_end, u_floor, k))
    a = oph.powerlaw_cool(u0, rho, 0.3 * t_end, u_floor, k)
    check("two half steps = one step (exact law)", oph.powerlaw_cool(a, rho, 0.3 * t_end, u_floor, k),
          oph.powerlaw_cool(u0, rho, 0.6 * t_end, u_floor, k), rtol=1e-12)  # fmt: skip


def test_cooling_units():
    """expected-values.md §7 (line 250): u(10^4 K) = 2.2961e-9 c^2 for the mu = 0.6 ideal gas."""
    check("u(1e4 K)", 2.2961e-9, oph.gas_u_from_temperature(1e4), rtol=1e-4)
    check("ideal floor = gas u at gamma 5/3", oph.gas_u_from_temperature(1e4), oph.ideal_u_floor(1e4, 5 / 3),
          rtol=1e-12)  # fmt: skip
    u_f = oph.ideal_u_floor(1e4, 5 / 3)
    check("beta cooling relaxes to the floor", u_f, oph.beta_cool(10 * u_f, u_f, 1e9, 1.0), rtol=1e-12)
