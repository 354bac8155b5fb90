"""Pinning the oracle's coupled step, its viscosity switch, velocity reconstruction, beta
cooling and time-step rule (CPU only).

These restate, for the FP64 oracle, the checks the reference holds for the same code:

- the shared step without the hole is plain kick-drift-kick (tests/sim/test_coupled.py:24-41);
- the hot-blob energy and angular-momentum ledgers through pericentre: max |dE| / sum m|H| below
  1e-5 at (eta_orb, C) = (0.00125, 0.15), the error falling 3-5x when both are halved, |dL|/|L|
  below 1e-10 (tests/sim/test_coupled.py:45-77), and the same ledger with beta cooling, whose
  radiated energy enters it (:80-93, here with the ideal-gas EOS);
- the Cullen-Dehnen switch formulas (tests/sim/test_shock.py:35-73) and a rigid rotation that
  leaves the switch off (:94-105);
- velocity reconstruction: a linear flow gets no viscosity, opposite gradients give the plain
  viscosity, the reconstructed force conserves momentum, angular momentum and energy
  (tests/sim/test_reconstruction.py:147-200), and the three gates act on a single pair as
  their definitions say (physics/viscosity.py:112-156);
- beta cooling: exact decay, radiated energy and the floor (tests/sim/test_cooling.py:17-59);
- stable_dt: the orbit criterion agrees with the reference's vectorised np.min(r**1.5) to two
  ulps, and a NaN in h / v_sig gives a NaN step as np.min does.

A wrong Strang split, a dropped kick, a wrong decay factor or a sign error in the reconstructed
jump fails one of them.
"""

import math

import numpy as np
import pytest

from helpers import T_END_BLOB, U_BLOB, check, hot_blob_config, hot_blob_ic, lattice_sphere

from oracle import coupled as oc
from oracle import neighbours as on
from oracle import physics as oph
from oracle import sph as osph


# ----------------------------------------------------------------- the step without the hole
def test_without_hole_equals_kick_drift_kick(rng):
    """hole=False: the coupled step is kick-drift-kick with the predicted velocity in the force
    pass (reference hydro.py:253-268), result for result (test_coupled.py:24-41 tolerances)."""
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.hydro import HydroConfig

    pos = lattice_sphere(700, 0.35)
    n = len(pos)
    v = rng.normal(0, 0.3, (n, 3))
    u = rng.uniform(0.5, 1.5, n)
    m = np.full(n, 1 / 4096)
    h0 = np.full(n, 1.6 * 4096 ** (-1 / 3))
    cfg = CoupledConfig(hydro=HydroConfig(), hole=False)
    st = oc.initialise(pos, v, u, m, h0, cfg)
    # the kick-drift-kick run, written from the reference's Phase-3 stepper; it keeps the
    # particles in the caller's order (no grid ordering), so sums run in another order
    x, vel, uu, h, alpha = pos.copy(), v.copy(), u.copy(), h0.copy(), None
    ev = oc._hydro(oc.State(0.0, x, vel, uu, m, h, np.full(n, cfg.hydro.alpha_init), perm=np.arange(n)),
                   cfg, vel, 0.0, None)  # fmt: skip
    acc, dudt, h, alpha = ev.a_hydro, ev.dudt, ev.h, ev.alpha
    dt = 0.05 * float(np.min(st.h / st.vsig_max))
    for _ in range(5):
        v_half = vel + 0.5 * dt * acc
        u_half = uu + 0.5 * dt * dudt
        x = x + dt * v_half
        v_pred = v_half + 0.5 * dt * acc
        pred = oc.State(0.0, x, v_pred, u_half + 0.5 * dt * dudt, m, h, alpha, perm=np.arange(n))
        ev = oc._hydro(pred, cfg, v_pred, dt, acc)
        acc, dudt, h, alpha = ev.a_hydro, ev.dudt, ev.h, ev.alpha
        vel = v_half + 0.5 * dt * acc
        uu = u_half + 0.5 * dt * dudt
        st = oc.step(st, cfg, dt)
    check("no hole: positions = kick-drift-kick", x, st.caller_order("x"), rtol=1e-14, per_vector=True)
    check("no hole: momenta = kick-drift-kick velocities", vel, st.caller_order("p"), rtol=1e-13, per_vector=True)
    check("no hole: internal energies equal", uu, st.caller_order("u"), rtol=1e-13)


# ----------------------------------------------------------------- hot-blob ledgers
def _hot_blob(eta_orb, courant, cooling_beta=math.inf):
    """blob.py:36-131 (shared steps): the hot blob from apocentre to T_END_BLOB on the oracle.
    Returns the largest ledger errors and what the blob did."""
    cfg = hot_blob_config(eta_orb, courant, cooling_beta=cooling_beta)
    st = oc.initialise(*hot_blob_ic(), cfg)
    led0 = oc.ledger(st, cfg)
    worst = {"e": 0.0, "l": 0.0, "u_max": 0.0, "l_min": np.inf, "steps": 0}
    while st.t < T_END_BLOB * (1 - 1e-14):
        st = oc.step(st, cfg, min(oc.stable_dt(st, cfg), T_END_BLOB - st.t))
        led = oc.ledger(st, cfg)
        worst["e"] = max(worst["e"], abs(led["energy"] - led0["energy"]) / led0["abs_h"])
        worst["l"] = max(worst["l"], float(np.linalg.norm(led["angmom"] - led0["angmom"]) / np.linalg.norm(led0["angmom"])))
        worst["u_max"] = max(worst["u_max"], float(st.u.max()))
        worst["l_min"] = min(worst["l_min"], float(np.min(np.linalg.norm(oph.gn_angmom(st.x, st.p), axis=1))))
        worst["steps"] += 1
    worst["e_radiated"] = st.e_radiated
    return worst


@pytest.fixture(scope="module")
def blob_runs():
    return {"fine": _hot_blob(0.00125, 0.15), "coarse": _hot_blob(0.0025, 0.3)}


def test_hot_blob_ledger(blob_runs):
    """test_coupled.py:45-77: (a) max ledger error < 1e-5 at (0.00125, 0.15); (b) it falls 3-5x
    from (0.0025, 0.3): the step is second order; (c) angular momentum to 1e-10; the blob is
    heated at pericentre and stays on safe orbits."""
    fine, coarse = blob_runs["fine"], blob_runs["coarse"]
    print(f"   hot blob: {coarse['steps']} / {fine['steps']} steps, ledger {coarse['e']:.3e} -> {fine['e']:.3e}")
    check("hot blob: max |dE| / sum m|H| (eta 0.00125, C 0.15)", 0.0, fine["e"], atol=1e-5)
    check("hot blob: error ratio coarse / fine", 4.0, coarse["e"] / fine["e"], atol=1.0)
    check("hot blob: max |dL| / |L|", 0.0, fine["l"], atol=1e-10)
    check("hot blob: heated at pericentre (u rose > 10x)", True, fine["u_max"] > 10 * U_BLOB)
    check("hot blob stays on safe orbits (lowest particle L > 5)", True, fine["l_min"] > 5.0)


def test_hot_blob_ledger_with_beta_cooling():
    """test_coupled.py:80-93 with the ideal-gas EOS: beta = 10 cooling in the middle of the
    split, its radiated energy in the ledger: max |d(sum m (H + u) + E_rad)| / sum m|H| < 1e-5,
    angular momentum to 1e-10, and energy is radiated. A wrong sign or a wrong radius in the
    cooling step's bookkeeping breaks the ledger."""
    fine = _hot_blob(0.00125, 0.15, cooling_beta=10.0)
    print(f"   hot blob, beta 10: ledger {fine['e']:.3e}, E_rad {fine['e_radiated']:.3e}")
    check("beta cooling: max ledger error (with E_rad)", 0.0, fine["e"], atol=1e-5)
    check("beta cooling: angular momentum", 0.0, fine["l"], atol=1e-10)
    check("beta cooling: energy radiated", True, fine["e_radiated"] > 0.0)


# ----------------------------------------------------------------- Cullen-Dehnen switch
def test_switch_formulas():
    """test_shock.py:35-73 (viscosity.py:82-106, Phantom Eqs. 46-50)."""
    check("Balsara xi: pure compression 1, pure rotation 0", (1.0, 0.0),
          (float(oph.balsara_xi(-1.0, 0.0)), float(oph.balsara_xi(0.0, 2.0))))  # fmt: skip
    check("Balsara xi: flow at rest 0", 0.0, float(oph.balsara_xi(0.0, 0.0)))
    check("Balsara xi: |div| = |curl| gives 1/2", 0.5, float(oph.balsara_xi(3.0, 3.0)), rtol=1e-15)
    check("alpha_loc capped at alpha_max", 1.0, float(oph.alpha_local(1.0, -1e9, 1.0, 1.0)))
    check("alpha_loc = 0 when compression is easing", 0.0, float(oph.alpha_local(1.0, +5.0, 1.0, 1.0)))
    # 10 h^2 xi max(-d div v / dt, 0) / c_s^2 below the cap: 10 * 0.01 * 0.5 * 2 / 4 = 0.025
    check("alpha_loc below the cap", 0.025, float(oph.alpha_local(0.1, -2.0, 0.5, 2.0)), rtol=1e-14)
    check("decay time h / (0.1 c_s)", 5.0, float(oph.decay_time(1.0, 2.0)), rtol=1e-15)
    check("alpha jumps up to alpha_loc at once", 0.8, float(oph.alpha_update(0.1, 0.8, 1e-3, 1.0)))
    tau, dt = 2.0, 0.5
    check("alpha decays by backward Euler", 1.0 / (1 + dt / tau), float(oph.alpha_update(1.0, 0.0, dt, tau)),
          rtol=1e-15)  # fmt: skip
    # backward Euler toward a nonzero alpha_loc: (alpha + a_loc dt/tau) / (1 + dt/tau)
    check("alpha decays toward alpha_loc", (0.9 + 0.2 * 0.25) / 1.25, float(oph.alpha_update(0.9, 0.2, dt, tau)),
          rtol=1e-15)  # fmt: skip


def _fcc(n_cells):
    a = 1.0 / n_cells
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    c = np.arange(n_cells)
    cells = np.stack(np.meshgrid(c, c, c, indexing="ij"), axis=-1).reshape(-1, 3)
    return (cells[:, None, :] + basis[None, :, :] + 0.25).reshape(-1, 3) * a


def test_rotation_does_not_trigger_switch():
    """test_shock.py:94-105: rigid rotation has no compression, so xi = 0 in the interior and
    alpha_loc = 0 there."""
    pos = _fcc(8) - 0.5
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 1.2 * (1.0 / n) ** (1 / 3)), kernel_name="cubic", eta=1.2)
    vel = np.cross([0, 0, 1.0], pos)
    g = osph.gradients_numba(pos, vel, np.cross([0, 0, 1.0], vel), mass, d.h, d.neighbours, "cubic")
    inner = np.linalg.norm(pos, axis=1) < 0.3
    xi = oph.balsara_xi(g.div_v[inner], g.curl_v[inner])
    check("rigid rotation: Balsara xi (interior)", 0.0, xi.max(), atol=1e-20)
    check("rigid rotation: |curl v| = 2 (interior)", np.full(int(inner.sum()), 2.0), g.curl_v[inner], rtol=1e-12)
    a_loc = oph.alpha_local(d.h[inner], g.ddivv_dt[inner], xi, np.full(int(inner.sum()), 0.3))
    check("rigid rotation: alpha_loc = 0 (interior)", 0.0, float(np.max(a_loc)), atol=0.0)


# ----------------------------------------------------------------- velocity reconstruction
def _smooth_sphere(n=2000, radius=0.45):
    pos = lattice_sphere(n, radius, seed=4)
    m = np.full(len(pos), 1 / 4096)
    d = osph.compute_density(pos, m, np.full(len(pos), 1.6 * 4096 ** (-1 / 3)), kernel_name="wendland_c4", eta=1.6)
    return pos, m, d


def _forces(pos, vel, m, d, p, **kw):
    return osph.forces_numba(pos, vel, m, d.h, d.rho, d.omega, p, d.neighbours, "wendland_c4", eta_crit=1 / 1.6,
                             **kw)  # fmt: skip


def test_linear_flow_has_no_viscosity():
    """test_reconstruction.py:147-174: a smooth compression (with shear and rotation on top)
    gets no artificial viscosity: the reconstructed jump v*_ab = (v_a - Phi/2 G_a x_ab) -
    (v_b + Phi/2 G_b x_ab) vanishes for a linear field. A wrong sign or a missing half in it
    leaves heating behind (shown with the negated gradient). Only x.G.x enters the projected
    jump and the limiter, so the antisymmetric part of G (hence its transpose) cannot matter."""
    pos, m, d = _smooth_sphere()
    n = len(pos)
    M = np.array([[-1.0, 0.6, -0.2], [-0.3, -0.8, 0.5], [0.4, -0.1, -1.2]])
    vel = pos @ M.T + np.array([0.1, -0.2, 0.05])
    u = np.full(n, 1e-4)
    p = (2 / 3) * d.rho * u
    cs = np.sqrt(10 / 9 * u)
    alpha = np.ones(n)
    gv = osph.velocity_gradients_numba(pos, vel, m, d.h, d.neighbours, "wendland_c4")
    check("linear field: corrected gradient exact", np.tile(M.ravel(), (n, 1)), gv.reshape(n, 9), rtol=1e-10,
          per_vector=True)  # fmt: skip
    base = _forces(pos, vel, m, d, p)
    plain = _forces(pos, vel, m, d, p, c_s=cs, alpha=alpha)
    recon = _forces(pos, vel, m, d, p, c_s=cs, alpha=alpha, reconstruction=True, closeness=False, grad_v=gv)
    close = _forces(pos, vel, m, d, p, c_s=cs, alpha=alpha, reconstruction=True, closeness=True, grad_v=gv)
    inner = np.linalg.norm(pos, axis=1) < 0.45 - 4 * np.median(d.h)
    heat = {k: float(np.abs(f.dudt - base.dudt)[inner].max()) for k, f in (("plain", plain), ("recon", recon),
                                                                           ("close", close))}  # fmt: skip
    print(f"   viscous heating, interior: {heat}")
    check("linear flow: reconstruction removes the viscous heating", 0.0, heat["recon"] / heat["plain"], atol=1e-10)
    check("linear flow: no viscous force either", base.acc[inner], recon.acc[inner], rtol=1e-10, per_vector=True)
    check("linear flow: with the closeness factor, heating < 10% of plain", True, heat["close"] < 0.1 * heat["plain"])
    # the negated gradient doubles the jump instead of removing it
    bad = _forces(pos, vel, m, d, p, c_s=cs, alpha=alpha, reconstruction=True, closeness=False,
                  grad_v=np.ascontiguousarray(-gv))  # fmt: skip
    check("a negated gradient leaves heating (the check can fail)", True,
          float(np.abs(bad.dudt - base.dudt)[inner].max()) > 1e-3 * heat["plain"])  # fmt: skip


def _random_state(rng, n=800):
    pos = lattice_sphere(n, 0.5, seed=9, jitter=0.3)
    m = rng.uniform(0.5, 1.5, n) / n
    d = osph.compute_density(pos, m, np.full(n, 1.6 * (1.0 / n) ** (1 / 3)), kernel_name="wendland_c4", eta=1.6)
    u = rng.uniform(0.5, 1.5, n)
    p = (2 / 3) * d.rho * u
    cs = np.sqrt(10 / 9 * u)
    vel = rng.normal(0, 0.5, pos.shape)
    alpha = rng.uniform(0, 1, n)
    return pos, vel, m, d, p, u, cs, alpha


def test_opposite_gradients_give_plain_viscosity(rng):
    """test_reconstruction.py:177-200: when every pair's two gradients disagree in sign
    (x.G_a.x and x.G_b.x of opposite signs, A < 0, Phi = 0), the force is the plain one."""
    pos, vel, m, d, p, u, cs, alpha = _random_state(rng)
    n = len(pos)
    gv = np.where((np.arange(n) % 2 == 0)[:, None, None], 1.0, -1.0) * np.eye(3)[None]
    nb = d.neighbours
    keep = [np.array([b for b in nb.of(a) if b == a or (a + b) % 2 == 1], np.int64) for a in range(n)]
    off = np.concatenate([[0], np.cumsum([len(k) for k in keep])]).astype(np.int64)
    pairs = on.Neighbours(off, np.concatenate(keep), nb.box)

    def f(**kw):
        return osph.forces_numba(pos, vel, m, d.h, d.rho, d.omega, p, pairs, "wendland_c4", u=u, c_s=cs,
                                 alpha=alpha, eta_crit=1 / 1.6, **kw)  # fmt: skip

    plain = f()
    for gate in (0, 2):
        recon = f(reconstruction=True, recon_gate=gate, grad_v=np.ascontiguousarray(gv))
        check(f"A < 0 for every pair (gate {gate}): acceleration = plain", plain.acc, recon.acc, rtol=1e-14,
              per_vector=True)  # fmt: skip
        check(f"A < 0 for every pair (gate {gate}): du/dt = plain", plain.dudt, recon.dudt, rtol=1e-13)


def test_reconstructed_forces_conserve(rng):
    """test_reconstruction.py:124-144: the reconstructed viscosity is a pair force, so sum m a,
    sum m x cross a and the total energy rate vanish to round-off."""
    pos, vel, m, d, p, u, cs, alpha = _random_state(rng)
    gv = osph.velocity_gradients_numba(pos, vel, m, d.h, d.neighbours, "wendland_c4")
    for gate in (0, 1, 2):
        f = osph.forces_numba(pos, vel, m, d.h, d.rho, d.omega, p, d.neighbours, "wendland_c4", u=u, c_s=cs,
                              alpha=alpha, reconstruction=True, recon_gate=gate, eta_crit=1 / 1.6,
                              grad_v=gv)  # fmt: skip
        rate = np.sum(m * (np.sum(vel * f.acc, axis=1) + f.dudt))
        scale = np.sum(m * (np.abs(np.sum(vel * f.acc, axis=1)) + np.abs(f.dudt)))
        check(f"recon gate {gate}: d(total energy)/dt (relative)", 0.0, abs(rate) / scale, atol=1e-13)
        mom = np.sum(m[:, None] * f.acc, axis=0)
        check(f"recon gate {gate}: sum m a (relative)", 0.0,
              np.linalg.norm(mom) / np.sum(m * np.linalg.norm(f.acc, axis=1)), atol=1e-13)  # fmt: skip
        am = np.sum(m[:, None] * np.cross(pos, f.acc), axis=0)
        am_scale = np.sum(m * np.linalg.norm(np.cross(pos, f.acc), axis=1))
        check(f"recon gate {gate}: sum m x cross a (relative)", 0.0, np.linalg.norm(am) / am_scale, atol=1e-13)


def test_gates_on_one_pair():
    """Two particles on the x axis, both with gradient G = g I (A = 1, Phi = 1, closeness off):
    the reconstructed jump is v*.r = dv - g r. Gate 1 (published) acts on v*.r alone, gate 0
    needs both v*.r and v.r negative, gate 2 (min-mod) takes the less compressive of the two
    (viscosity.py:151-156). The viscous heating of particle a is proportional to
    (alpha c_s + beta |mu|) mu (v.r), so ratios of it between cases are fixed numbers."""
    r = 0.1
    pos = np.array([[r, 0.0, 0.0], [0.0, 0.0, 0.0]])
    m = np.ones(2)
    h = np.full(2, 0.12)
    nb = on.Neighbours(np.array([0, 2, 4], np.int64), np.array([0, 1, 0, 1], np.int64), None)
    rho, omega, p = np.ones(2), np.ones(2), np.zeros(2)
    cs, alpha = np.full(2, 0.5), np.ones(2)

    def heat(dv, gr, gate):
        vel = np.array([[dv, 0.0, 0.0], [0.0, 0.0, 0.0]])
        gv = np.ascontiguousarray(np.broadcast_to((gr / r) * np.eye(3), (2, 3, 3)))
        f = osph.forces_numba(pos, vel, m, h, rho, omega, p, nb, "cubic", c_s=cs, alpha=alpha, reconstruction=True,
                              closeness=False, recon_gate=gate, grad_v=gv)  # fmt: skip
        return f.dudt_visc[0]

    def expect(mu, vdotr):
        return (alpha[0] * cs[0] + 2.0 * abs(mu)) * mu * vdotr

    # receding pair (v.r = +0.1) whose reconstructed jump approaches (v*.r = 0.1 - 0.3 = -0.2)
    check("receding, v*.r < 0: gate 0 off", 0.0, heat(0.1, 0.3, 0))
    check("receding, v*.r < 0: gate 2 off", 0.0, heat(0.1, 0.3, 2))
    check("receding, v*.r < 0: gate 1 acts (and cools)", True, heat(0.1, 0.3, 1) < 0.0)
    # approaching pair whose reconstructed jump recedes: every gate off
    for gate in (0, 1, 2):
        check(f"approaching, v*.r > 0: gate {gate} off", 0.0, heat(-0.1, -0.3, gate))
    # both approach, v.r = -0.2 > v*.r = -0.3: min-mod uses -0.2, gates 0 and 1 use -0.3
    ref = heat(-0.2, 0.1, 0)
    check("both approach: gate 1 = gate 0", ref, heat(-0.2, 0.1, 1), rtol=1e-15)
    check("both approach: min-mod / gate 0 = mu ratio", expect(-0.2, -0.2) / expect(-0.3, -0.2),
          heat(-0.2, 0.1, 2) / ref, rtol=1e-13)  # fmt: skip
    # against the plain viscosity of the same pair (mu = v.r): the ratio of the mu factors
    plain = osph.forces_numba(pos, np.array([[-0.2, 0.0, 0.0], [0.0, 0.0, 0.0]]), m, h, rho, omega, p, nb, "cubic",
                              c_s=cs, alpha=alpha).dudt_visc[0]  # fmt: skip
    check("both approach: gate 0 / plain = mu ratio", expect(-0.3, -0.2) / expect(-0.2, -0.2), ref / plain,
          rtol=1e-13)  # fmt: skip
    check("plain heating of an approaching pair is positive", True, plain > 0.0)


# ----------------------------------------------------------------- beta cooling
def test_beta_cooling_exact_decay_and_ledger():
    """test_cooling.py:17-32: u(t) = u_floor + (u0 - u_floor) exp(-t/t_cool) to 1e-12;
    t_cool = beta / Omega_K = beta r^1.5; m (u0 - u1) is the radiated energy."""
    r, beta, u0, uf, m = 100.0, 10.0, 1e-5, 2.3e-9, 0.3
    tc = float(oph.cooling_time(r, beta))
    check("t_cool = beta r^1.5", beta * r**1.5, tc, rtol=1e-15)
    for t in (0.1 * tc, tc, 3.7 * tc):
        check(f"u at t = {t / tc:.1f} t_cool", uf + (u0 - uf) * math.exp(-t / tc), float(oph.beta_cool(u0, uf, t, tc)),
              rtol=1e-12)  # fmt: skip
    u1 = float(oph.beta_cool(u0, uf, tc, tc))
    check("radiated energy m (u0 - u1)", m * (u0 - uf) * (1 - math.exp(-1)), m * (u0 - u1), rtol=1e-12)
    check("beta = inf: u unchanged", u0, float(oph.beta_cool(u0, uf, tc, math.inf)))


def test_beta_cooling_floor():
    """test_cooling.py:35-50: after 100 t_cool u sits at the floor, from above and from below,
    never crossing it on the way down; the floor is the ideal-gas energy at 10^4 K."""
    uf, tc = 2.3e-9, 5.0
    check("from above: u = u_floor after 100 t_cool", uf, float(oph.beta_cool(1e-4, uf, 100 * tc, tc)), rtol=1e-15)
    check("from below: u = u_floor", uf, float(oph.beta_cool(1e-10, uf, 100 * tc, tc)), rtol=1e-15)
    path = oph.beta_cool(1e-4, uf, np.linspace(0, 100 * tc, 1000), tc)
    check("never below the floor on the way down", True, bool(np.all(path >= uf)))
    # T(u_floor) = 10^4 K for the ideal gas: T = (gamma - 1) mu m_p u c^2 / k_B
    gamma = 5.0 / 3.0
    u_f = oph.ideal_u_floor(1e4, gamma)
    check("ideal gas: T(u_floor) = 1e4 K", 1e4, (gamma - 1.0) * oph.MU * oph.M_P * u_f * oph.C**2 / oph.K_B,
          rtol=1e-12)  # fmt: skip


# ----------------------------------------------------------------- stable_dt
def _dt_state(rng, n=300):
    x = rng.normal(0, 1, (n, 3)) * 50.0 + np.array([3000.0, 0.0, 0.0])
    st = oc.State(0.0, x, np.zeros((n, 3)), np.ones(n), np.ones(n), rng.uniform(1, 2, n), np.ones(n),
                  a_hydro=rng.normal(0, 1e-6, (n, 3)), vsig_max=rng.uniform(1e-3, 2e-3, n))  # fmt: skip
    return st


def test_stable_dt_orbit_term_matches_the_reference_form(rng):
    """coupled.py:213-226 takes eta_orb * np.min(r**1.5) (NumPy's vectorised power); the oracle
    takes eta_orb * min(r)**1.5 (the scalar pow the device path also uses). pow is monotone, so
    both pick the same particle; the two powers differ by at most an ulp, the product by two."""
    from nebula.sim.coupled import baseline_config

    cfg = baseline_config()
    for _ in range(20):
        st = _dt_state(rng)
        st.a_hydro[:] = 0.0  # the orbit criterion decides
        st.vsig_max[:] = 1e-12
        r = np.linalg.norm(st.x, axis=1)
        ref = cfg.eta_orb * np.min(r**1.5)
        got = oc.stable_dt(st, cfg)
        check("orbit criterion: oracle vs np.min(r**1.5)", ref, got, rtol=4.5e-16)


def test_stable_dt_nan_propagates_like_np_min(rng):
    """np.min propagates NaN: a NaN in h / v_sig gives a NaN step in the reference. The device
    minima propagate NaN the same way, and dt_from_minima (the host half of the device
    stable_dt) keeps Python's min(dt, x) semantics: NaN minima of the other criteria are
    dropped, a NaN Courant minimum makes dt NaN."""
    from nebula.sim.coupled import baseline_config, dt_from_minima

    cfg = baseline_config()
    st = _dt_state(rng)
    st.h[7] = np.nan
    check("oracle: NaN h gives a NaN step", True, math.isnan(oc.stable_dt(st, cfg)))
    check("device host half: NaN Courant minimum gives NaN", True,
          math.isnan(dt_from_minima(np.array([np.nan, 1.0, np.inf, 3000.0]), cfg, False)))  # fmt: skip
    ok = dt_from_minima(np.array([2.0, np.nan, np.inf, np.nan]), cfg, False)
    check("device host half: NaN force / orbit minima are dropped (min(dt, nan) = dt)", cfg.hydro.courant * 2.0, ok)
    st2 = _dt_state(rng)
    st2.a_hydro[3] = np.nan  # excluded by |a| > 0, as in the reference
    h_ref = st2.h.copy()
    amag = np.linalg.norm(st2.a_hydro, axis=1)
    keep = amag > 0
    expect = min(cfg.hydro.courant * np.min(h_ref / st2.vsig_max), cfg.hydro.c_force * np.min(np.sqrt(h_ref[keep] / amag[keep])),
                 cfg.eta_orb * np.min(np.linalg.norm(st2.x, axis=1) ** 1.5))  # fmt: skip
    check("oracle: a NaN acceleration is left out", expect, oc.stable_dt(st2, cfg), rtol=2.3e-16)
