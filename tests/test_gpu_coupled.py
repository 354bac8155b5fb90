"""GPU parity of the whole coupled step (coupled.initialise / stable_dt / step / ledger) against
the FP64 oracle on identical seeded inputs: the state after several steps is bit-identical
(every operation of the step is the oracle's own sequence of roundings); the ledger's sums
(whose summation order differs) agree to 1e-13 of sum m|H|."""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _run(ic, cfg, steps):
    from nebula.sim import coupled

    from oracle import coupled as oc

    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for k in range(steps):
        dt = coupled.stable_dt(st, cfg)
        odt = oc.stable_dt(ost, cfg)
        check(f"step {k}: stable_dt bit-identical", odt, dt)
        st = coupled.step(st, cfg, dt)
        ost = oc.step(ost, cfg, odt)
    return st, ost


def _compare(st, ost, cfg, fields=("x", "p", "u", "h", "rho", "alpha", "a_hydro", "dudt", "vsig_max", "c_s")):
    from nebula.sim import coupled

    from oracle import coupled as oc

    got = st.to_numpy()
    bit_equal("state.perm (particle order of the grid)", ost.perm, st.perm.cpu().numpy())
    for k in fields:
        bit_equal(f"state.{k}", ost.caller_order(k), got[k])
    lg, lo = coupled.ledger(st, cfg), oc.ledger(ost, cfg)
    scale = lo["abs_h"]
    check("ledger energy (1e-13 of sum m|H|)", lo["energy"] / scale, lg["energy"] / scale, atol=1e-13)
    check("ledger e_radiated (1e-13 of sum m|H|)", lo["e_radiated"] / scale, lg["e_radiated"] / scale, atol=1e-13)
    ls = float(np.sum(ost.mass * np.linalg.norm(np.cross(ost.x, ost.p), axis=1)))
    check("ledger angular momentum (1e-13 of sum m|x cross p|)", lo["angmom"] / ls, lg["angmom"] / ls, atol=1e-13)


def test_cloud_on_parabola_plain_fp64():
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    ic = jittered_sphere_on_parabola(2048, seed=20261014)
    cfg = coupled.baseline_config()
    print("coupled: 2048-particle cloud, plain FP64, 4 steps", flush=True)
    st, ost = _run(ic, cfg, 4)
    _compare(st, ost, cfg)


def test_stream_crash_cooling_balsara(monkeypatch):
    """Past the first step (whose smoothing lengths grow from the initial conditions and rebuild
    the grid) the FP64 step's density pass is the lane-group one (nb_hydro_density64): the
    comparison with the oracle covers it."""
    from dataclasses import replace

    from nebula import _hip
    from nebula.ics import stream_crash
    from nebula.sim import coupled

    called = []
    real = _hip.call

    def spy(name, *a):
        called.append(name)
        return real(name, *a)

    monkeypatch.setattr(_hip, "call", spy)
    ic = stream_crash(8192, length=200.0, seed=20261014)
    base = coupled.baseline_config(balsara=True)
    cfg = replace(base, cooling_powerlaw=True)
    print("coupled: 8192-particle stream crash, power-law cooling + Balsara, 4 steps", flush=True)
    st, ost = _run(ic, cfg, 4)
    lane_passes = called.count("nb_hydro_tight64")
    print("density passes on the lane-group lists:", lane_passes, "of", called.count("nb_hydro_tight") + lane_passes)
    check("steps after the first took the lane-group density pass", True, lane_passes >= 3)
    _compare(st, ost, cfg)


def test_beta_cooling_step():
    """The reference's in-step beta cooling (coupled.py:244-251): t_cool = beta r_mid^1.5 at the
    mid-step radius, exact relaxation toward the 10^4 K floor, E_rad += sum m (u_half - u_cool).
    The relaxation factor exp(-dt / t_cool) and r^-1.5 come from the device's libm, which may
    differ from NumPy's in the last bit, so the state is compared at 1e-12 relative (the FP64
    mode's bound) rather than bitwise; E_rad (a device sum against the host's pairwise sum) to
    1e-12 of itself."""
    from dataclasses import replace

    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    ic = jittered_sphere_on_parabola(2048, r0=3000.0, r_p=500.0, seed=5)
    ic = replace(ic, u=ic.u * 100.0)  # 10^6 K: well above the 10^4 K floor the cooling relaxes to
    cfg = replace(coupled.baseline_config(), cooling_beta=10.0)
    print("coupled: 2048-particle cloud, beta = 10 cooling, 4 steps", flush=True)
    from nebula.sim import coupled as c2

    from oracle import coupled as oc

    st = c2.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for _ in range(4):
        dt = oc.stable_dt(ost, cfg)
        check("stable_dt (1e-12 relative)", dt, c2.stable_dt(st, cfg), rtol=1e-12)
        st = c2.step(st, cfg, dt)
        ost = oc.step(ost, cfg, dt)
    got = st.to_numpy()
    for k in ("x", "p"):
        check(f"beta cooling: {k} (1e-12 per vector)", ost.caller_order(k), got[k], rtol=1e-12, per_vector=True)
    for k in ("u", "h", "rho", "vsig_max"):
        check(f"beta cooling: {k} (1e-12 relative)", ost.caller_order(k), got[k], rtol=1e-12)
    for k in ("a_hydro", "dudt"):  # sums that cancel: also 1e-12 of their rms
        ref = ost.caller_order(k)
        rms = float(np.sqrt(np.mean(ref**2) * (3 if ref.ndim > 1 else 1)))
        check(f"beta cooling: {k} (1e-12 relative + 1e-12 rms)", ref, got[k], rtol=1e-12, atol=1e-12 * rms,
              per_vector=ref.ndim > 1)  # fmt: skip
    same = np.mean(ost.caller_order("u") == got["u"])
    print(f"u bit-identical for {same:.4f} of the particles")
    check("radiated energy (1e-12 relative)", ost.e_radiated, float(st.e_radiated), rtol=1e-12)
    check("energy was radiated", True, ost.e_radiated > 0.0)


def test_switch_conductivity_reconstruction():
    """The reference's default hole physics: Cullen-Dehnen switch, velocity-signal conductivity,
    velocity reconstruction (exp() in the closeness factor: compared at 1e-12)."""
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    ic = jittered_sphere_on_parabola(1024, r0=3000.0, r_p=200.0, seed=7)
    cfg = coupled.CoupledConfig(hydro=coupled.HydroConfig(kernel="cubic", eta=1.2, h_max=200.0,
                                                          conductivity_signal=1, alpha_u=0.05))  # fmt: skip
    from nebula.sim import coupled as c2

    from oracle import coupled as oc

    st = c2.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for _ in range(3):
        dt = oc.stable_dt(ost, cfg)
        st = c2.step(st, cfg, dt)
        ost = oc.step(ost, cfg, dt)
    got = st.to_numpy()
    check("switch+cond+recon: x", ost.caller_order("x"), got["x"], rtol=1e-12, per_vector=True)
    check("switch+cond+recon: a_hydro", ost.caller_order("a_hydro"), got["a_hydro"], rtol=1e-9, per_vector=True)
    check("switch+cond+recon: rho", ost.caller_order("rho"), got["rho"], rtol=1e-12)
    check("switch+cond+recon: alpha", ost.caller_order("alpha"), got["alpha"], rtol=1e-9, atol=1e-15)


def test_blocked_lists_grid_order_and_tighten(rng):
    """The step's own list layout (64-row blocks) holds the reference's lists; the grid order
    equals the oracle's restatement of it; in grid order no row needs sorting."""
    import torch

    from nebula import _dev, _hip
    from nebula.ics import stream_crash
    from nebula.sim import neighbours as nbm

    from oracle import neighbours as on

    ic = stream_crash(20000, length=400.0, seed=3)
    hs = 1.4 * ic.h * rng.uniform(0.9, 1.1, len(ic.h))
    x_d, h_d = _dev.f64(ic.x), _dev.f64(hs)
    order = nbm.grid_order(x_d, h_d)
    bit_equal("grid order = oracle restatement", on.grid_order(ic.x, hs), order.cpu().numpy())
    xo, ho = ic.x[order.cpu().numpy()], hs[order.cpu().numpy()]
    ref = on.find_neighbours_grid(xo, ho)
    blk = nbm.build_lists_blocked(_dev.f64(xo), _dev.f64(ho), None)
    stats = np.zeros(4, np.int64)
    _hip.check(_hip.lib().nb_lists_stats(_hip.ctx(), stats.ctypes.data_as(_hip._P64)))
    check("rows needing a sort in grid order", 0, int(stats[2]))
    csr = blk.to_csr()
    bit_equal("blocked lists: offsets", ref.offset, csr.offset)
    bit_equal("blocked lists: indices", ref.index, csr.index)
    t_ref = on.tighten(ref, xo, 0.8 * ho)
    t_blk = nbm.tighten_blocked(blk, _dev.f64(xo), _dev.f64(0.8 * ho)).to_csr()
    bit_equal("blocked tighten: offsets", t_ref.offset, t_blk.offset)
    bit_equal("blocked tighten: indices", t_ref.index, t_blk.index)
    # caller order: rows need sorting, still identical lists
    ref2 = on.find_neighbours_grid(ic.x, hs)
    blk2 = nbm.build_lists_blocked(x_d, h_d, None).to_csr()
    bit_equal("blocked lists (caller order): indices", ref2.index, blk2.index)
    del torch


@pytest.mark.parametrize("masses", ["uniform", "per-particle"])
def test_fp64_step_lists_in_blocks_of_their_own_capacity(masses):
    """The FP64 step's density pass sizes its 64-row blocks from the previous pass's row counts
    (two per particle), gathers again at exact capacities when a row outgrows its block, and
    writes the tight rows into blocks as long as their longest merged row. Nothing depends on the
    hint (none, the pass's own counts, a hint far too small): outputs and lists are bit-identical,
    and the tight rows are the reference's lists at the final h (neighbours.py:240-281: r_ij <
    2 max(h_i, h_j), which a search at h_search >= h holds in full)."""
    import torch

    from nebula import _dev
    from nebula.ics import stream_crash
    from nebula.sim import coupled, sph

    from oracle import neighbours as on

    ic = stream_crash(20000, length=400.0, seed=5)
    n = len(ic.h)
    mass = ic.mass if masses == "uniform" else ic.mass * np.random.default_rng(2).uniform(0.5, 1.5, n)
    hc = coupled.baseline_config().hydro
    x_d, m_d = _dev.f64(ic.x), _dev.f64(mass)
    mu = _dev.uniform_mass(m_d)

    def run(hint):
        counts = _dev.empty((n, 2), torch.int32)
        d, _ = sph.hydro_density_device(x_d, m_d, _dev.f64(ic.h), hc.kernel, hc.eta, 1e-4, hc.h_max, hc.skin, mu,
                                        "fp64", row_hint=hint, row_counts=counts)  # fmt: skip
        return d, counts

    print(f"fp64 blocks ({masses} masses): no hint, own counts, hint too small", flush=True)
    d0, c0 = run(None)
    d1, c1 = run(c0)
    d2, c2 = run(torch.zeros_like(c0))
    for name, d, c in (("own counts", d1, c1), ("hint too small", d2, c2)):
        for k in ("h", "rho", "omega", "iterations", "flags"):
            bit_equal(f"{name}: {k}", getattr(d0, k).cpu().numpy(), getattr(d, k).cpu().numpy())
        bit_equal(f"{name}: row counts", c0.cpu().numpy(), c.cpu().numpy())
        t0, t = d0.neighbours.to_csr(), d.neighbours.to_csr()
        bit_equal(f"{name}: tight offsets", t0.offset, t.offset)
        bit_equal(f"{name}: tight indices", t0.index, t.index)
    ref = on.find_neighbours_grid(ic.x, d0.h.cpu().numpy())
    t0 = d0.neighbours.to_csr()
    bit_equal("tight rows = reference lists at the final h: offsets", ref.offset, t0.offset)
    bit_equal("tight rows = reference lists at the final h: indices", ref.index, t0.index)
    boff = d0.neighbours.block_offset.cpu().numpy()
    ln = d0.neighbours.length.cpu().numpy()
    cap = np.diff(boff) // 64
    check("block starts are multiples of 64 apart", True, bool(np.all(np.diff(boff) % 64 == 0)))
    check("every tight row fits its block", True, bool(np.all(ln <= np.repeat(cap, 64)[:n])))
    check("merged rows (gather + extras) hold the tight rows", True, bool(np.all(c0.cpu().numpy().sum(1) >= ln)))
    print(f"tight entries allocated {boff[-1]} vs uniform blocks {((n + 63) // 64) * 64 * int(cap.max())}")


@pytest.mark.parametrize("case", ["cubic, uniform masses", "wendland, per-particle masses", "eta 2, light particles"])
def test_fp64_lane_lists_equal_per_thread_pass(case, monkeypatch):
    """nb_hydro_density64 / nb_hydro_tight64 (16-lane gather over trimmed column runs, FP64 terms
    added in row order, tight rows merged from the gather row and the sorted extras) against
    nb_hydro_density / nb_hydro_tight on particles stored in grid order: h, rho, omega,
    iterations, flags, row counts, the blocks' offsets, the rows' lengths and every tight row
    bit-identical, and the tight rows the reference's lists at the final h. The third case (eta
    2, Wendland: about 270 neighbours a row; 64 particles of 1e-4 the mass, whose smoothing
    lengths come out a fraction of their neighbours') has rows of more than 128 extras, which one
    lane merges."""
    import torch

    from nebula import _dev, _hip
    from nebula.ics import stream_crash
    from nebula.sim import coupled, sph
    from nebula.sim.neighbours import grid_order

    from oracle import neighbours as on

    light = case.startswith("eta 2")
    kernel = "cubic" if case.startswith("cubic") else "wendland_c4"
    eta = 2.0 if light else (1.2 if kernel == "cubic" else 1.6)
    ic = stream_crash(20000, length=400.0, eta=eta, seed=11)
    n = len(ic.h)
    rng = np.random.default_rng(4)
    mass, h0 = ic.mass.copy(), ic.h.copy()
    if case.startswith("wendland"):
        mass = mass * rng.uniform(0.5, 1.5, n)
    if light:
        few = rng.choice(n, 64, replace=False)
        mass[few] *= 1e-4
    hc = coupled.baseline_config().hydro
    x_d, m_d, h_d = _dev.f64(ic.x), _dev.f64(mass), _dev.f64(h0)
    mu = _dev.uniform_mass(m_d)
    # start from solved smoothing lengths (as the step does after its first): one build, whose
    # grid is the one the particles are ordered by
    h_d = sph.hydro_density_device(x_d, m_d, h_d, kernel, eta, 1e-4, hc.h_max, hc.skin, mu, "fp64")[0].h
    hs = _dev.empty(n)
    _hip.call("nb_search_lengths", _hip.ptr(h_d), float(hc.h_max), float(hc.skin), float(np.inf), n, _hip.ptr(hs))
    order = grid_order(x_d, hs).long()
    x_d, m_d, h_d = x_d[order].contiguous(), m_d[order].contiguous(), h_d[order].contiguous()

    called = []
    real = _hip.call

    def spy(name, *a):
        called.append(name)
        return real(name, *a)

    monkeypatch.setattr(_hip, "call", spy)

    def run(lanes):
        monkeypatch.setattr(sph, "FP64_LANE_LISTS", lanes)
        called.clear()
        counts = _dev.empty((n, 2), torch.int32)
        d, n_search = sph.hydro_density_device(x_d, m_d, h_d.clone(), kernel, eta, 1e-4, hc.h_max, hc.skin, mu,
                                               "fp64", row_counts=counts)  # fmt: skip
        return d, counts, n_search, list(called)

    print(f"fp64 lane-group lists ({case}, {n} particles in grid order)", flush=True)
    d0, c0, s0, calls0 = run(False)
    d1, c1, s1, calls1 = run(True)
    print("builds:", d0.builds, d1.builds, "longest extras segment:", int(c1[:, 1].max()), "calls:", calls1)
    check("the per-thread pass ran without the lane-group entries", False, "nb_hydro_density64" in calls0)
    check("the final lists came from the lane-group pass", True, "nb_hydro_tight64" in calls1)
    for k in ("h", "rho", "omega", "iterations", "flags"):
        bit_equal(f"{case}: {k}", getattr(d0, k).cpu().numpy(), getattr(d1, k).cpu().numpy())
    bit_equal(f"{case}: row counts", c0.cpu().numpy(), c1.cpu().numpy())
    check(f"{case}: search-list entries", s0, s1)
    bit_equal(f"{case}: block offsets", d0.neighbours.block_offset.cpu().numpy(),
              d1.neighbours.block_offset.cpu().numpy())  # fmt: skip
    bit_equal(f"{case}: row lengths", d0.neighbours.length.cpu().numpy(), d1.neighbours.length.cpu().numpy())
    t0, t1 = d0.neighbours.to_csr(), d1.neighbours.to_csr()
    bit_equal(f"{case}: tight offsets", t0.offset, t1.offset)
    bit_equal(f"{case}: tight indices", t0.index, t1.index)
    ref = on.find_neighbours_grid(x_d.cpu().numpy(), d1.h.cpu().numpy())
    bit_equal(f"{case}: tight rows = reference lists at the final h: offsets", ref.offset, t1.offset)
    bit_equal(f"{case}: tight rows = reference lists at the final h: indices", ref.index, t1.index)
    if light:
        check("rows of more than 128 extras were merged", True, int(c1[:, 1].max()) > 128)


def test_torus_shakura_sunyaev_cooling():
    """The torus workload's physics: Shakura-Sunyaev viscosity (alpha 0.1, stress form, which
    adds the N-S gradient and force passes) with power-law cooling, against the oracle; then two
    steps of the same in mixed mode with the learned start (finite).

    nu = alpha c_s^2 r^1.5 takes a non-integer power. The GPU rounds r^1.5 once (correctly
    rounded); the reference's NumPy pow (glibc, within 0.52 ulp) differs from that by one ulp for
    about 1 value in 1000, so this comparison is at 1e-12 relative (the FP64 mode's bound), not
    bitwise: x, p, u, h, rho per value (acc and du/dt per vector over their rms)."""
    from dataclasses import replace

    from nebula.ics import torus
    from nebula.sim import coupled
    from nebula.sim.learned_h import LearnedH

    ic = torus(4096, seed=5)
    cfg = replace(coupled.baseline_config(ns_alpha=0.1), cooling_powerlaw=True)
    print("coupled: 4096-particle torus, Shakura-Sunyaev alpha 0.1 + power-law cooling, 3 steps", flush=True)
    from oracle import coupled as oc

    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for k in range(3):
        dt, odt = coupled.stable_dt(st, cfg), oc.stable_dt(ost, cfg)
        check(f"step {k}: stable_dt (1e-12 relative)", odt, dt, rtol=1e-12)
        st = coupled.step(st, cfg, odt)
        ost = oc.step(ost, cfg, odt)
    got = st.to_numpy()
    for k in ("x", "p", "u", "h", "rho", "nu"):
        check(f"state.{k} (1e-12 relative)", ost.caller_order(k), got[k], rtol=1e-12, per_vector=k in ("x", "p"))
    for k in ("a_hydro", "dudt"):
        ref = ost.caller_order(k)
        rms = float(np.sqrt(np.mean(np.sum(ref.reshape(len(ref), -1) ** 2, axis=1))))
        check(f"state.{k} (1e-12 of rms)", ref, got[k], atol=1e-12 * rms, per_vector=k == "a_hydro")
    model = LearnedH(seed=1)
    cm = replace(cfg, hydro=replace(cfg.hydro, precision="mixed", learned_h=True, learned_model=model))
    print("coupled: the same torus, mixed + learned start, 2 steps", flush=True)
    sm = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cm)
    for _ in range(2):
        sm = coupled.step(sm, cm, coupled.stable_dt(sm, cm))
    got = sm.to_numpy()
    check("mixed torus: all fields finite", True,
          all(bool(np.all(np.isfinite(got[k]))) for k in ("x", "p", "u", "h", "rho", "a_hydro", "dudt", "nu")))
