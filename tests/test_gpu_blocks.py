"""Block (individual power-of-two) time steps on the GPU (nebula.sim.blocks) against the oracle's
restatement and the reference's tests/sim/test_blocks.py.

- Row subsets of the device passes equal the full passes row for row; list reuse with an active
  subset rebuilds when an inactive h outgrows the lists (test_blocks.py:36-109).
- Mixed levels with the limiter and wake-up (hot spots in a warm ball on the hot-blob orbit), and
  forced level offsets without the hole: every field of the state, the levels, the step ends,
  the wake-up count and the active set of every system step bit-identical to the oracle.
- Beta cooling in the middle of each particle's step against the oracle (1e-12: the device's exp
  and pow against NumPy's).
- One level without the limiter equals the GPU shared step (test_blocks.py:115-140; 1e-12, the
  shared step visiting particles in grid order).
- Gravity only, two orbits sharing the timeline: expected-values §2 pericentre time, radial
  period and precession (test_blocks.py:147-195, the second orbit circular).
- The hot-blob ledger on block steps (test_blocks.py:320-337): max |dE| / sum m|H| < 1e-5 at
  (eta_orb, C) = (0.00125, 0.15), falling 3-5x from (0.0025, 0.3); |dL|/|L| falls >= 3x.
"""

from dataclasses import replace

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check, hot_blob_config, hot_blob_ic, lattice_sphere

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FIELDS = ("x", "p", "u", "h", "rho", "omega", "alpha", "a_hydro", "dudt", "vsig_max", "level", "n_beg", "n_end",
          "x_beg", "p_half", "u_half", "heat_visc", "heat_cond", "heat_disk", "heat_pdv")  # fmt: skip


def test_active_rows_identical_to_full_pass(rng):
    from nebula.sim.neighbours import tighten
    from nebula.sim.sph import compute_density, density_numba, forces_numba, gradients_numba

    pos = lattice_sphere(900, 0.4)
    n = len(pos)
    m = np.full(n, 1 / 4096)
    vel = rng.normal(0, 0.3, (n, 3))
    acc = rng.normal(0, 1.0, (n, 3))
    h0 = np.full(n, 1.6 * 4096 ** (-1 / 3))
    rows = np.sort(rng.choice(n, n // 3, replace=False))
    print("row subsets: 900 particles, a third active", flush=True)
    # from converged h, as in a block step (the row subset then takes the same rebuild rounds)
    h0 = compute_density(pos, m, h0).h
    full = compute_density(pos, m, h0)
    part = compute_density(pos, m, h0, rows=rows)
    check("same rebuild rounds", full.builds, part.builds)
    for f in ("h", "rho", "omega"):
        bit_equal(f"density rows: {f}", getattr(full, f)[rows], getattr(part, f))
    lists = full.neighbours
    d2 = density_numba(pos, m, h0, lists, np.full(n, np.inf), rows=rows[::-1])
    d1 = density_numba(pos, m, h0, lists, np.full(n, np.inf))
    bit_equal("density_numba rows (reversed order)", d1.rho[rows[::-1]], d2.rho)
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
    bit_equal("forces rows: acceleration", f_all.acc[rows], f_row.acc)
    bit_equal("forces rows: du/dt", f_all.dudt[rows], f_row.dudt)
    bit_equal("forces rows: v_sig,max", f_all.vsig_max[rows], f_row.vsig_max)
    g_all = gradients_numba(pos, vel, acc, m, h, nb_all)
    g_row = gradients_numba(pos, vel, acc, m, h, nb_rows, rows=rows)
    bit_equal("gradients rows: div v", g_all.div_v[rows], g_row.div_v)
    bit_equal("gradients rows: d(div v)/dt", g_all.ddivv_dt[rows], g_row.ddivv_dt)


def test_reused_lists_stay_complete_for_inactive_neighbours(rng):
    from nebula.sim.neighbours import find_neighbours_numba
    from nebula.sim.sph import compute_density

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
    full = find_neighbours_numba(moved, h_big)
    complete = all(set(full.of(i)) <= set(res.neighbours.of(i)) for i in rows)
    check("lists rebuilt (not reused) when inactive h outgrows them", first.cache.builds + 1, res.cache.builds)
    check("no reuse counted", 0, res.cache.reuses)
    check("active rows' lists contain every neighbour", True, complete)


def _warm(seed=3, hot=True):
    from test_oracle_blocks import warm_ball

    rng = np.random.default_rng(seed)
    x, v, u, m, h = warm_ball(rng)
    if hot:
        u = u * np.where(np.arange(len(u)) % 5 == 0, 30.0, 1.0)
    return x, v, u, m, h


def _compare(st, ost, tag, rtol=None):
    got = st.to_numpy()
    for f in FIELDS:
        ref = getattr(ost, f)
        if rtol is None or f in ("level", "n_beg", "n_end"):
            bit_equal(f"{tag}: {f}", ref, got[f])
        else:
            check(f"{tag}: {f} ({rtol:g} relative)", ref, got[f], rtol=rtol, atol=1e-300, per_vector=ref.ndim == 2)


@pytest.mark.parametrize("case", ["hole_hot_spots", "hole_hot_spots_recon", "no_hole_offsets"])
def test_blocks_match_oracle(case):
    """Bit for bit without velocity reconstruction; with it to 1e-12 (the FP64 force pass's
    reconstruction branch matches the oracle's to 1e-12, not bitwise)."""
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig

    from oracle import blocks as ob
    from test_oracle_blocks import warm_ball

    cc = CoupledConfig()
    recon = case.endswith("recon")
    cc = replace(cc, hydro=replace(cc.hydro, reconstruction=recon))
    if case.startswith("hole_hot_spots"):
        x, v, u, m, h = _warm()
        cfg = BlockConfig(coupled=cc, dt_max=64.0)
        offsets = None
    else:
        x, v, u, m, h = warm_ball(np.random.default_rng(5))
        x = x - np.array([500.0, 0.0, 0.0])
        v = v - np.array([0.0, 7.1951756691 / 500.0, 0.0])
        cfg = BlockConfig(coupled=replace(cc, hole=False), dt_max=2.0)
        offsets = np.zeros(len(x), np.int64)
        offsets[np.random.default_rng(7).choice(len(x), len(x) // 2, replace=False)] = 2
    print(f"blocks vs oracle: {case}, {len(x)} particles, 40 system steps", flush=True)
    st = blocks.initialise(x, v, u, m, h, cfg, t_stop=cfg.dt_max * 8, level_offset=offsets)
    ost = ob.initialise(x, v, u, m, h, cfg, t_stop=cfg.dt_max * 8, level_offset=offsets)
    tol = 1e-12 if recon else None
    _compare(st, ost, "initialise", tol)
    for k in range(40):
        rows = blocks.advance(st, cfg)
        orows = ob.advance(ost, cfg)
        bit_equal(f"system step {k}: active set", orows, rows)
        check(f"system step {k}: tick", ost.n, st.n)
    _compare(st, ost, "after 40 system steps", tol)
    check("wake-ups", ost.woken, st.woken)
    levels = np.unique(st.to_numpy()["level"])
    print(f"levels in use {levels.tolist()}, woken {st.woken}, active counts {st.active_counts[:12]}...")
    check("mixed levels in use", True, len(levels) >= 2)


def test_beta_cooling_matches_oracle():
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig

    from oracle import blocks as ob

    x, v, u, m, h = _warm()
    cfg = BlockConfig(coupled=CoupledConfig(cooling_beta=10.0), dt_max=64.0)
    print("blocks with beta = 10 cooling vs oracle, 30 system steps", flush=True)
    st = blocks.initialise(x, v, u * 100.0, m, h, cfg, t_stop=512.0)
    ost = ob.initialise(x, v, u * 100.0, m, h, cfg, t_stop=512.0)
    for _ in range(30):
        bit_equal("active set", ob.advance(ost, cfg), blocks.advance(st, cfg))
    _compare(st, ost, "beta cooling", rtol=1e-12)
    # each booked term m (u_half - u_cooled) cancels: an ulp of the device exp against NumPy's shows
    # at the scale of the thermal energy, which bounds it
    e_th = float(np.sum(m * u * 100.0))
    check("radiated energy (1e-12 of the thermal energy)", ost.e_radiated, st.e_radiated, atol=1e-12 * e_th)
    check("energy was radiated", True, ost.e_radiated > 0)
    check("heat_cool (1e-12 of u)", ost.heat_cool, st.to_numpy()["heat_cool"], atol=1e-12 * float(np.max(u * 100.0)))


@pytest.mark.parametrize("ns_nu, disc_alpha_av", [(0.0, 0.0), (0.02, 0.0), (0.0, 1.0)])
def test_one_level_equals_shared_step(ns_nu, disc_alpha_av):
    from nebula.sim import blocks, coupled
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.hydro import HydroConfig
    from nebula.sim.sph import COND_VELOCITY

    x, v, u, m, h = _warm(seed=4, hot=False)
    hc = HydroConfig(conductivity_signal=COND_VELOCITY, ns_nu=ns_nu, disc_alpha_av=disc_alpha_av)
    cc = CoupledConfig(hydro=hc)
    cfg = BlockConfig(coupled=cc, dt_max=1.0, limiter=False)
    print(f"one level vs shared step: ns_nu {ns_nu}, disc alpha {disc_alpha_av}", flush=True)
    st = blocks.initialise(x, v, u, m, h, cfg, t_stop=6.0)
    ref = coupled.initialise(x, v, u, m, h, cc)
    check("all particles on level 0", 0, int(st.to_numpy()["level"].max()))
    for _ in range(6):
        st.neighbour_cache = None  # the shared step builds its lists afresh at every pass
        blocks.advance(st, cfg)
        ref = coupled.step(ref, cc, 1.0)
    got, want = st.to_numpy(), ref.to_numpy()
    check("one level: positions", want["x"], got["x"], rtol=1e-12, per_vector=True)
    check("one level: momenta", want["p"], got["p"], rtol=1e-12, per_vector=True)
    check("one level: internal energies", want["u"], got["u"], rtol=1e-12)
    check("one level: time", ref.t, st.t, rtol=0.0)
    if ns_nu > 0 or disc_alpha_av > 0:
        check("one level: disk viscosity heated every particle", True, bool(np.all(got["heat_disk"] > 0)))


def test_orbits_under_block_steps():
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig

    from oracle import physics as ph
    from oracle.integrator import gn_v_from_p
    from test_oracle_blocks import circular_state, pericentres
    from test_oracle_gravity import apocentre_state

    xc, _, pc = apocentre_state(25.0, 5000.0)
    xb, pb = circular_state(2000.0)
    x0, p0 = np.array([xc, xb]), np.array([pc, pb])
    cfg = BlockConfig(coupled=CoupledConfig(eta_orb=0.00125), dt_max=4096.0, hydro=False)
    print("gravity only: the cloud orbit and a circular one on the block timeline", flush=True)
    st = blocks.initialise(x0, gn_v_from_p(x0, p0), np.zeros(2), np.ones(2), np.ones(2), cfg, 310 * 4096.0)
    rec = {0: [(0.0, xc, pc)], 1: [(0.0, xb, pb)]}
    while st.n < st.n_stop:
        rows = blocks.advance(st, cfg)
        xs, ps = st.x.cpu().numpy(), st.p.cpu().numpy()
        for i in rows:
            rec[int(i)].append((st.t, xs[i].copy(), ps[i].copy()))
    t = np.array([r[0] for r in rec[0]])
    x = np.array([r[1] for r in rec[0]])
    t_peri, phi_peri = pericentres(t, x)
    check("cloud pericentres found", 2, len(t_peri))
    check("§2 first pericentre time (block steps)", 396_223.627, t_peri[0], rtol=1e-5)
    check("§2 radial period (block steps)", 792_447.254, t_peri[1] - t_peri[0], rtol=1e-5)
    check("§2 precession (block steps, deg)", 24.0398, np.degrees(np.diff(phi_peri)[0] - 2 * np.pi), atol=0.01)
    for i in (0, 1):
        xi = np.array([r[1] for r in rec[i]])
        pi = np.array([r[2] for r in rec[i]])
        hh = ph.gn_hamiltonian(xi, pi)
        ll = np.linalg.norm(ph.gn_angmom(xi, pi), axis=1)
        check(f"orbit {i}: max |dH/H|", 0.0, float(np.max(np.abs(hh / hh[0] - 1))), atol=1e-4)
        check(f"orbit {i}: max |dL/L|", 0.0, float(np.max(np.abs(ll / ll[0] - 1))), atol=1e-12)
    print(f"system steps {len(st.active_counts)}, particle steps {st.particle_steps}")


def test_hot_blob_ledger_block_steps():
    """test_blocks.py:320-337 (blob.py:98-110: dt_max 64, ledger at every full sync point)."""
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig

    from helpers import T_END_BLOB

    out = {}
    for label, eta, cour in (("fine", 0.00125, 0.15), ("coarse", 0.0025, 0.3)):
        cc = hot_blob_config(eta, cour)
        cfg = BlockConfig(coupled=cc, dt_max=64.0)
        print(f"hot blob on block steps ({label}: eta_orb {eta}, C {cour})", flush=True)
        st = blocks.initialise(*hot_blob_ic(), cfg, t_stop=T_END_BLOB)
        led0 = blocks.ledger(st, cfg)
        e = lm = 0.0
        while st.n < st.n_stop:
            blocks.advance(st, cfg)
            if not st.synced():
                continue
            led = blocks.ledger(st, cfg)
            e = max(e, abs(led["energy"] - led0["energy"]) / led0["abs_h"])
            lm = max(lm, float(np.linalg.norm(led["angmom"] - led0["angmom"]) / np.linalg.norm(led0["angmom"])))
        out[label] = (e, lm, st.particle_steps, len(st.active_counts))
        print(f"  max |dE|/sum m|H| {e:.3e}, |dL|/|L| {lm:.3e}, particle-steps {st.particle_steps}, "
              f"system steps {len(st.active_counts)}, levels {np.unique(st.to_numpy()['level']).tolist()}",
              flush=True)  # fmt: skip
    fe, ce = out["fine"][0], out["coarse"][0]
    fl, cl = out["fine"][1], out["coarse"][1]
    check("block steps: hot blob max |dE| / sum m|H|", 0.0, fe, atol=1e-5)
    check("block steps: error ratio coarse / fine", 4.0, ce / fe, atol=1.0)
    check("block steps: |dL|/|L| falls >= 3x per halving", True, cl / fl >= 3.0)


def test_synchronization_and_coverage():
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig
    from nebula.sim.timesteps import ticks

    x, v, u, m, h = _warm()
    cfg = BlockConfig(coupled=CoupledConfig(), dt_max=64.0)
    st = blocks.initialise(x, v, u, m, h, cfg, t_stop=64.0 * 8)
    covered = np.zeros(len(x), np.int64)
    beg = st.to_numpy()["n_beg"]
    ok_active = ok_grid = ok_inside = True
    for _ in range(80):
        before_end = st.to_numpy()["n_end"]
        rows = blocks.advance(st, cfg)
        a = st.to_numpy()
        ok_active &= bool(np.all(before_end[rows] == st.n))
        covered[rows] += st.n - beg[rows]
        beg[rows] = st.n
        lv = a["level"]
        ok_grid &= bool(np.all(a["n_beg"] % ticks(lv) == 0) and np.all(a["n_end"] % ticks(lv) == 0))
        ok_inside &= bool(np.all((a["n_beg"] <= st.n) & (st.n < a["n_end"])))
    check("active particles exactly at their step end", True, ok_active)
    check("step boundaries on each particle's level grid", True, ok_grid)
    check("no particle past its step end", True, ok_inside)
    check("completed steps tile the timeline (no gaps)", st.to_numpy()["n_beg"], covered)
    d = blocks.diagnostics(st)
    check("diagnostics: particle-steps", int(sum(st.active_counts)), d["particle_steps"])


def test_rejected_settings():
    from nebula.sim import blocks
    from nebula.sim.blocks import BlockConfig
    from nebula.sim.coupled import CoupledConfig

    x, v, u, m, h = _warm()
    for cfg in (BlockConfig(r_acc=4.0), BlockConfig(respa=True),
                BlockConfig(coupled=replace(CoupledConfig(), cooling_powerlaw=True))):  # fmt: skip
        with pytest.raises(NotImplementedError):
            blocks.initialise(x, v, u, m, h, cfg, t_stop=64.0)
