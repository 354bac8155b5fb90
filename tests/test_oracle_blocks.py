"""Pinning the oracle's block time steps (CPU only; the reference's tests/sim/test_blocks.py
restated on the oracle, plus the level arithmetic of timesteps.py).

- Level arithmetic: level_for_dt is the smallest k with dt_max / 2^k <= dt (brute force), the
  rule for longer steps, the limiter's fixed point (neighbours within max_gap, and minimal: no
  raised level can come down one).
- The SPH passes restricted to a subset of rows equal the full passes row for row, and list reuse
  with an active subset rebuilds when an inactive h outgrows the lists (test_blocks.py:36-109).
- Every particle on level 0 without the limiter is the shared-step scheme, step for step, with
  Navier-Stokes and disc viscosity too (test_blocks.py:115-140; to 1e-12 here, because the
  oracle's shared step visits particles in grid order and the block scheme in the caller's).
- Gravity only, two orbits sharing the timeline: pericentre time, radial period and precession of
  expected-values §2 (test 1, test_blocks.py:147-195).
- Synchronisation and coverage of the timeline with mixed levels, limiter and wake-up
  (test_blocks.py:265-289).
"""

import numpy as np
import pytest

from helpers import check, lattice_sphere
from test_oracle_gravity import apocentre_state

from oracle import blocks as ob
from oracle import coupled as oc
from oracle import timesteps as ots
from oracle.integrator import gn_v_from_p
from oracle.neighbours import find_neighbours_grid, tighten
from oracle.sph import compute_density, density_numba, forces_numba, gradients_numba


def warm_ball(rng, u=1e-3, sigma_v=0.01, n=505):
    """test_blocks.py:24-31: a ball of radius 3 at x = (500, 0, 0) on the hot-blob orbit, warmer
    than the blob and stirred (a jittered lattice in place of the reference's glass)."""
    pos = lattice_sphere(n, 3.0) + np.array([500.0, 0.0, 0.0])
    rho = 1.0 / (4 / 3 * np.pi * 27.0)
    v = np.tile([0.0, 7.1951756691 / 500.0, 0.0], (n, 1)) + rng.normal(0, sigma_v, (n, 3))
    return pos, v, np.full(n, u), np.full(n, 1.0 / n), np.full(n, 1.6 * (1 / n / rho) ** (1 / 3))


def block_config(**kw):
    from nebula.sim.blocks import BlockConfig

    return BlockConfig(**kw)


# ----------------------------------------------------------------- level arithmetic
def test_level_for_dt_is_the_largest_power_of_two_step(rng):
    dt_max = 4096.0
    dt = np.concatenate([10.0 ** rng.uniform(-6, 4, 4000), dt_max / 2.0 ** np.arange(0, 30),
                         np.nextafter(dt_max / 2.0 ** np.arange(0, 30), 0), np.nextafter(dt_max / 2.0 ** np.arange(0, 30), 1e9)])  # fmt: skip
    k = ots.level_for_dt(dt, dt_max)
    brute = np.array([next(j for j in range(41) if dt_max / 2.0**j <= d) for d in dt])
    check("level_for_dt = smallest k with dt_max 2^-k <= dt (brute force)", brute, k)
    with pytest.raises(ValueError):
        ots.level_for_dt([0.0], dt_max)
    with pytest.raises(ValueError):
        ots.level_for_dt([dt_max / 2.0**41], dt_max)


def test_allowed_level_rules():
    t = ots.ticks
    check("ticks(K_MAX) = 1, ticks(0) = 2^40", [1, 2**40], t([40, 0]).tolist())
    # shorter: always; longer: only to a level whose grid n lies on
    n = 3 * t(5)  # a multiple of level 5's step, not of level 4's
    check("shorter step always allowed", [7], ots.allowed_level(n, [5], [7]).tolist())
    check("longer step: the longest level n is on", [5], ots.allowed_level(n, [6], [2]).tolist())
    check("longer step: n on level 3's grid", [3], ots.allowed_level(4 * t(3), [6], [3]).tolist())
    check("is_active", [True, False], ots.is_active(n, [5, 4]).tolist())


def test_limiter_fixed_point(rng):
    pos = rng.uniform(0, 10, (600, 3))
    nb = find_neighbours_grid(pos, np.full(600, 0.6))
    lv = rng.integers(0, 4, 600)
    lv[rng.choice(600, 5, replace=False)] = 12  # sharp jumps propagate outward
    out = ots.limit_levels(lv, nb, 2)
    owner = np.repeat(np.arange(600), np.diff(nb.offset))
    check("levels only rise", True, bool(np.all(out >= lv)))
    check("neighbours within 2 levels", True, bool(np.all(out[owner] >= out[nb.index] - 2)))
    raised = np.flatnonzero(out > lv)
    minimal = all(np.any(out[nb.index[nb.offset[i] : nb.offset[i + 1]]] - 2 > out[i] - 1) for i in raised)
    check("minimal: no raised level can come down one", True, minimal)
    check("the jump spread", True, len(raised) > 20)


# ----------------------------------------------------------------- active subsets
def test_active_rows_identical_to_full_pass(rng):
    """test_blocks.py:36-83"""
    pos = lattice_sphere(900, 0.4)
    n = len(pos)
    m = np.full(n, 1 / 4096)
    vel = rng.normal(0, 0.3, (n, 3))
    acc = rng.normal(0, 1.0, (n, 3))
    h0 = np.full(n, 1.6 * 4096 ** (-1 / 3))
    rows = np.sort(rng.choice(n, n // 3, replace=False))
    # start from converged h, as a block step does: a row subset then goes through the same
    # rebuild rounds as the full pass (from a poor start the full pass can take a round more,
    # for particles outside the subset, which re-solves every row from its first answer)
    h0 = compute_density(pos, m, h0).h
    full = compute_density(pos, m, h0)
    part = compute_density(pos, m, h0, rows=rows)
    check("same rebuild rounds", full.builds, part.builds)
    for f in ("h", "rho", "omega"):
        check(f"density rows: {f}", getattr(full, f)[rows], getattr(part, f), rtol=0.0)
    lists = full.neighbours
    d2 = density_numba(pos, m, h0, lists, np.full(n, np.inf), rows=rows[::-1])
    d1 = density_numba(pos, m, h0, lists, np.full(n, np.inf))
    check("density_numba rows (reversed order)", d1.rho[rows[::-1]], d2.rho, rtol=0.0)
    h, rho, omega = full.h, full.rho, full.omega
    u = rng.uniform(0.5, 1.5, n)
    p_gas = (2 / 3) * rho * u
    cs = np.sqrt(10 / 9 * u)
    alpha = rng.uniform(0, 1, n)
    nb_all = tighten(lists, pos, h)
    nb_rows = tighten(lists, pos, h, rows=rows)
    same = all(np.array_equal(nb_all.of(i), nb_rows.of(i)) for i in rows)
    others = np.setdiff1d(np.arange(n), rows)
    check("tighten rows: same lists for rows, empty for the rest", (True, True),
          (same, bool(np.all(np.diff(nb_rows.offset)[others] == 0))))  # fmt: skip
    f_all = forces_numba(pos, vel, m, h, rho, omega, p_gas, nb_all, u=u, c_s=cs, alpha=alpha)
    f_row = forces_numba(pos, vel, m, h, rho, omega, p_gas, nb_rows, u=u, c_s=cs, alpha=alpha, rows=rows)
    check("forces rows: acceleration", f_all.acc[rows], f_row.acc, rtol=0.0)
    check("forces rows: du/dt", f_all.dudt[rows], f_row.dudt, rtol=0.0)
    check("forces rows: v_sig,max", f_all.vsig_max[rows], f_row.vsig_max, rtol=0.0)
    g_all = gradients_numba(pos, vel, acc, m, h, nb_all)
    g_row = gradients_numba(pos, vel, acc, m, h, nb_rows, rows=rows)
    check("gradients rows: div v", g_all.div_v[rows], g_row.div_v, rtol=0.0)
    check("gradients rows: d(div v)/dt", g_all.ddivv_dt[rows], g_row.ddivv_dt, rtol=0.0)


def test_reused_lists_stay_complete_for_inactive_neighbours(rng):
    """test_blocks.py:86-109"""
    pos = lattice_sphere(900, 0.4)
    n = len(pos)
    m = np.full(n, 1 / 4096)
    h0 = np.full(n, 1.6 * 4096 ** (-1 / 3))
    first = compute_density(pos, m, h0)
    rows = np.arange(0, n, 2)
    h_big = first.h.copy()
    h_big[1::2] *= 1.5
    moved = pos + rng.normal(0, 1e-4, pos.shape)
    res = compute_density(moved, m, h_big, cache=first.cache, rows=rows)
    full = find_neighbours_grid(moved, h_big)
    complete = all(set(full.of(i)) <= set(res.neighbours.of(i)) for i in rows)
    check("lists rebuilt (not reused) when inactive h outgrows them", first.cache.builds + 1, res.cache.builds)
    check("no reuse counted", 0, res.cache.reuses)
    check("active rows' lists contain every neighbour", True, complete)


# ----------------------------------------------------------------- one level
@pytest.mark.parametrize("ns_nu, disc_alpha_av", [(0.0, 0.0), (0.02, 0.0), (0.0, 1.0)])
def test_one_level_equals_shared_step(rng, ns_nu, disc_alpha_av):
    """test_blocks.py:115-140. The shared step here builds its lists afresh at every pass, so
    the block run drops its list cache before each system step (with reuse the Newton solve of
    h starts from a different search bracket and lands elsewhere within its 1e-4 tolerance;
    test_reuse_moves_h_within_the_newton_tolerance)."""
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.hydro import HydroConfig
    from nebula.sim.sph import COND_VELOCITY

    x, v, u, m, h = warm_ball(rng)
    dt = 1.0
    hc = HydroConfig(conductivity_signal=COND_VELOCITY, ns_nu=ns_nu, disc_alpha_av=disc_alpha_av)
    cc = CoupledConfig(hydro=hc)
    cfg = block_config(coupled=cc, dt_max=dt, limiter=False)
    st = ob.initialise(x, v, u, m, h, cfg, t_stop=6 * dt)
    ref = oc.initialise(x, v, u, m, h, cc)
    check("all particles on level 0", 0, int(np.max(st.level)))
    for _ in range(6):
        st.neighbour_cache = None
        ob.advance(st, cfg)
        ref = oc.step(ref, cc, dt)
    check("one level: positions", ref.caller_order("x"), st.x, rtol=1e-12, per_vector=True)
    check("one level: momenta", ref.caller_order("p"), st.p, rtol=1e-12, per_vector=True)
    check("one level: internal energies", ref.caller_order("u"), st.u, rtol=1e-12)
    check("one level: viscosity alphas", ref.caller_order("alpha"), st.alpha, rtol=1e-10, atol=1e-14)
    check("one level: time", ref.t, st.t, rtol=0.0)
    if ns_nu > 0 or disc_alpha_av > 0:
        check("one level: disk viscosity heated every particle", True, bool(np.all(st.heat_disk > 0)))


def test_reuse_moves_h_within_the_newton_tolerance(rng):
    """The same one-level run with the list cache kept: h, rho and the state stay within what the
    Newton tolerance (1e-4 of h per solve) allows of the run without reuse."""
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.hydro import HydroConfig
    from nebula.sim.sph import COND_VELOCITY

    x, v, u, m, h = warm_ball(rng)
    cc = CoupledConfig(hydro=HydroConfig(conductivity_signal=COND_VELOCITY))
    cfg = block_config(coupled=cc, dt_max=1.0, limiter=False)
    a = ob.initialise(x, v, u, m, h, cfg, t_stop=6.0)
    b = ob.initialise(x, v, u, m, h, cfg, t_stop=6.0)
    for _ in range(6):
        a.neighbour_cache = None
        ob.advance(a, cfg)
        ob.advance(b, cfg)
    print(f"reuses {b.neighbour_cache.reuses}, builds {b.neighbour_cache.builds}")
    check("lists were reused", True, b.neighbour_cache.reuses >= 3)
    check("h within 1e-4", a.h, b.h, rtol=1e-4)
    check("rho within 3e-4", a.rho, b.rho, rtol=3e-4)
    check("u within 1e-6", a.u, b.u, rtol=1e-6)


# ----------------------------------------------------------------- test 1
def pericentres(t, x):
    """orbit_metrics.py:12-30: local minima of r refined by parabolas through three samples."""
    r = np.linalg.norm(x, axis=1)
    phi = np.unwrap(np.arctan2(x[:, 1], x[:, 0]))
    idx = np.where((r[1:-1] < r[:-2]) & (r[1:-1] <= r[2:]))[0] + 1
    times, angles = [], []
    for i in idx:
        sl = slice(i - 1, i + 2)
        c = np.polyfit(t[sl], r[sl], 2)
        tp = -c[1] / (2 * c[0])
        times.append(tp)
        angles.append(np.polyval(np.polyfit(t[sl] - tp, phi[sl], 2), 0.0))
    return np.array(times), np.array(angles)


def circular_state(r):
    """A circular GN orbit of radius r in the x-y plane, on the -x axis."""
    vt = (r - 2) / (r * np.sqrt(r - 3))
    x = np.array([-r, 0.0, 0.0])
    return x, ph_p_from_v(x, np.array([0.0, -vt, 0.0]))


def ph_p_from_v(x, v):
    from oracle import physics as ph

    return ph.gn_p_from_v(x[None], v[None])[0]


def test_orbits_under_block_steps():
    """test_blocks.py:147-195 with the second orbit made circular (r = 2000) so that the oracle's
    Python loop stays short: hydro off, the cloud orbit (r_p 25, r_a 5,000) and the circular one
    sharing the block timeline at eta = 0.00125, to 1.6 cloud periods: expected-values §2
    pericentre time and radial period (1e-5), precession (0.01 deg), max |dH/H| < 1e-4 and
    |dL/L| < 1e-12 (the GPU test runs the reference's hot-blob orbit as the second one)."""
    from nebula.sim.coupled import CoupledConfig

    from oracle import physics as ph

    xc, _, pc = apocentre_state(25.0, 5000.0)
    xb, pb = circular_state(2000.0)
    x0, p0 = np.array([xc, xb]), np.array([pc, pb])
    cfg = block_config(coupled=CoupledConfig(eta_orb=0.00125), dt_max=4096.0, hydro=False)
    t_end = 310 * 4096.0
    st = ob.initialise(x0, gn_v_from_p(x0, p0), np.zeros(2), np.ones(2), np.ones(2), cfg, t_end)
    rec = {0: [(0.0, xc, pc)], 1: [(0.0, xb, pb)]}
    while st.n < st.n_stop:
        for i in ob.advance(st, cfg):
            rec[int(i)].append((st.t, st.x[i].copy(), st.p[i].copy()))
    t = np.array([r[0] for r in rec[0]])
    x = np.array([r[1] for r in rec[0]])
    t_peri, phi_peri = pericentres(t, x)
    check("block steps: cloud pericentres found", 2, len(t_peri))
    check("§2 first pericentre time (block steps)", 396_223.627, t_peri[0], rtol=1e-5)
    check("§2 radial period (block steps)", 792_447.254, t_peri[1] - t_peri[0], rtol=1e-5)
    check("§2 precession (block steps, deg)", 24.0398, np.degrees(np.diff(phi_peri)[0] - 2 * np.pi), atol=0.01)
    for i, name in ((0, "cloud"), (1, "circular orbit")):
        xi = np.array([r[1] for r in rec[i]])
        pi = np.array([r[2] for r in rec[i]])
        hh = ph.gn_hamiltonian(xi, pi)
        ll = np.linalg.norm(ph.gn_angmom(xi, pi), axis=1)
        check(f"{name}: max |dH/H|", 0.0, float(np.max(np.abs(hh / hh[0] - 1))), atol=1e-4)
        check(f"{name}: max |dL/L|", 0.0, float(np.max(np.abs(ll / ll[0] - 1))), atol=1e-12)
    print(f"system steps {len(st.active_counts)}, particle steps {st.particle_steps}")


# ----------------------------------------------------------------- test 6
def mixed_run(n_adv=80, seed=3):
    rng = np.random.default_rng(seed)
    x, v, u, m, h = warm_ball(rng)
    u = u * np.where(np.arange(len(u)) % 5 == 0, 30.0, 1.0)  # hot spots: mixed levels
    from nebula.sim.coupled import CoupledConfig

    cfg = block_config(coupled=CoupledConfig(), dt_max=64.0)
    return (x, v, u, m, h), cfg


def test_synchronization_and_coverage():
    """test_blocks.py:265-289"""
    (x, v, u, m, h), cfg = mixed_run()
    st = ob.initialise(x, v, u, m, h, cfg, t_stop=64.0 * 8)
    covered = np.zeros(len(x), np.int64)
    beg = st.n_beg.copy()
    ok_active = ok_grid = ok_inside = True
    for _ in range(80):
        before_end = st.n_end.copy()
        rows = ob.advance(st, cfg)
        ok_active &= bool(np.all(before_end[rows] == st.n))
        covered[rows] += st.n - beg[rows]
        beg[rows] = st.n
        lv = st.level
        ok_grid &= bool(np.all(st.n_beg % ots.ticks(lv) == 0) and np.all(st.n_end % ots.ticks(lv) == 0))
        ok_inside &= bool(np.all((st.n_beg <= st.n) & (st.n < st.n_end)))
    print(f"levels in use {np.unique(st.level).tolist()}, woken {st.woken}, system steps {len(st.active_counts)}")
    check("levels in use (mixed)", True, len(np.unique(st.level)) >= 2)
    check("active particles exactly at their step end", True, ok_active)
    check("step boundaries on each particle's level grid", True, ok_grid)
    check("no particle past its step end", True, ok_inside)
    check("completed steps tile the timeline (no gaps)", st.n_beg, covered)
