"""GPU tests of the mixed-precision mode (FP32 pair terms, compensated sums, FP64 where the
method needs it) against the FP64 oracle on identical seeded inputs.

Neighbour lists stay exact in this mode: the search predicate is evaluated in FP64, and the
tightened lists are the oracle's tighten() of the search lists at the mixed pass's own h.
Everything else carries FP32 rounding, bounded per field as follows (measured errors are
printed; the bounds leave about a factor ten of room):

  density pass   h, rho          relative 1e-5        omega       absolute 1e-5
  forces         acc             |da| <= 1e-5 rms|a|  dudt        |dd| <= 1e-5 rms|dudt|
                 vsig_max        relative 1e-6
  gradients      div v, |curl v| |d| <= 1e-4 rms|div v|, d(div v)/dt  |d| <= 1e-4 rms
  coupled step   x               1e-6 of the particle's h after 4 steps
                 u               relative 1e-4;  h, rho relative 1e-5;  a_hydro 1e-4 rms|a|
                 energy ledger   1e-9 of sum m|H| (the ledger is dominated by the orbital energy)
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _rms(a):
    a = np.asarray(a, float)
    return float(np.sqrt(np.mean(np.sum(a.reshape(len(a), -1) ** 2, axis=1))))


def _scaled(name, ref, got, bound):
    """max |got - ref| (per vector for (N,3)) over the rms size of ref must stay below bound."""
    ref = np.asarray(ref, float)
    got = np.asarray(got, float)
    d = got - ref
    err = np.linalg.norm(d, axis=1) if d.ndim == 2 else np.abs(d)
    val = float(err.max()) / _rms(ref)
    ok = val <= bound and not np.any(np.isnan(got))
    print(("PASS " if ok else "FAIL ") + f"{name}: max error / rms = {val:.3g} (bound {bound:g})")
    assert ok, name


def _relative(name, ref, got, bound):
    ref = np.asarray(ref, float)
    got = np.asarray(got, float)
    val = float(np.max(np.abs(got - ref) / np.abs(ref)))
    ok = val <= bound and not np.any(np.isnan(got))
    print(("PASS " if ok else "FAIL ") + f"{name}: max relative error {val:.3g} (bound {bound:g})")
    assert ok, name


def _grid_ordered_stream(n, seed):
    from nebula.ics import stream_crash

    from oracle import neighbours as on

    ic = stream_crash(n, length=400.0, seed=seed)
    hs = np.minimum(ic.h, 200.0) * 1.4
    o = on.grid_order(ic.x, hs)
    return ic.x[o], ic.mass[o], ic.h[o]


@pytest.mark.parametrize("masses", ["uniform", "varied"])
def test_mixed_density_pass(masses):
    from nebula import _dev
    from nebula.sim.sph import hydro_density_mixed_device

    from oracle import neighbours as on
    from oracle import sph as osph

    x, m, h0 = _grid_ordered_stream(20000, 5)
    if masses == "uniform":
        m = np.full(len(m), float(m.mean()))
    m_uniform = float(m[0]) if masses == "uniform" else 0.0
    print(f"mixed density pass: stream crash n={len(x)}, {masses} masses", flush=True)
    ref = osph.compute_density(x, m, h0, kernel_name="cubic", eta=1.2, h_max=200.0, skin=1.4)
    runs = []
    for _ in range(2):
        d, n_search = hydro_density_mixed_device(_dev.f64(x), _dev.f64(m), _dev.f64(h0, copy=True), "cubic", 1.2,
                                                 1e-4, 200.0, 1.4, m_uniform)  # fmt: skip
        runs.append((d, n_search))
    d, n_search = runs[0]
    check("rebuild rounds", ref.builds, d.builds)
    h = _dev.host(d.h)
    _relative("h", ref.h, h, 1e-5)
    _relative("rho", ref.rho, _dev.host(d.rho), 1e-5)
    check("omega (absolute 1e-5)", ref.omega, _dev.host(d.omega), atol=1e-5)
    check("flags", ref.flags, _dev.host(d.flags))
    # the lists are exact: search lists of the oracle, tightened at the mixed h
    check("search-list entries", int(ref.neighbours.offset[-1]), n_search)
    t_ref = on.tighten(ref.neighbours, x, h)
    t_got = d.neighbours.to_csr()
    bit_equal("tight lists (mixed h): offsets", t_ref.offset, t_got.offset)
    bit_equal("tight lists (mixed h): indices", t_ref.index, t_got.index)
    check("tight-list total", int(t_ref.offset[-1]), d.neighbours.n_total)
    # run to run the same bits (extras are put in index order before the sums)
    d2 = runs[1][0]
    bit_equal("repeat: rho", _dev.host(d.rho), _dev.host(d2.rho))
    bit_equal("repeat: tight rows (entries in emission order)", _row_entries(d.neighbours), _row_entries(d2.neighbours))


def _row_entries(nb):
    """The entries of every row, in stored order, concatenated (the gaps between rows dropped)."""
    ln = nb.length.cpu().numpy().astype(np.int64)
    off = nb.offset.cpu().numpy()
    idx = nb.index.cpu().numpy()
    take = np.repeat(off[:-1] - np.concatenate([[0], np.cumsum(ln)[:-1]]), ln) + np.arange(int(ln.sum()))
    return idx[take]


@pytest.mark.parametrize("mode", ["pressure", "visc", "visc_cond_p", "visc_cond_v"])
def test_mixed_forces(rng, mode):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import ShockCapturing, forces_numba
    from test_gpu_kernels import _moving_state

    from oracle import sph as osph

    pos, vel, mass, d, p, u, cs, alpha, nb = _moving_state(rng)
    use_visc = mode != "pressure"
    use_u = mode.startswith("visc_cond")
    sig = 1 if mode.endswith("_v") else 0
    kw = dict(u=u if use_u else None, c_s=cs if use_visc else None, alpha=alpha if use_visc else None)
    print(f"mixed forces: mode={mode} n={len(pos)}", flush=True)
    ref = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", beta=2.0, alpha_u=0.05,
                            cond_signal=sig, **kw)  # fmt: skip
    shock = ShockCapturing(beta=2.0, alpha_u=0.05, conductivity_signal=sig)
    got = forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, Neighbours(nb.offset, nb.index, None), "cubic",
                       shock=shock, precision="mixed", **kw)  # fmt: skip
    _scaled(f"{mode}: acc", ref.acc, got.acc, 1e-5)
    _scaled(f"{mode}: dudt", ref.dudt, got.dudt, 1e-5)
    if use_visc:
        _relative(f"{mode}: vsig_max", ref.vsig_max, got.vsig_max, 1e-6)
        _scaled(f"{mode}: dudt_visc", ref.dudt_visc, got.dudt_visc, 1e-5)
    if use_u:
        _scaled(f"{mode}: dudt_cond", ref.dudt_cond, got.dudt_cond, 1e-5)


def test_mixed_forces_periodic_and_rows(rng):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import ShockCapturing, forces_numba

    from oracle import sph as osph
    from oracle.neighbours import tighten

    n, box = 3000, 1.0
    pos = rng.uniform(0, box, (n, 3))
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 0.05), box=box, kernel_name="cubic", eta=1.2)
    nb = tighten(d.neighbours, pos, d.h)
    vel = rng.normal(0, 0.2, (n, 3))
    u = rng.uniform(0.5, 1.5, n)
    p = (2.0 / 3.0) * d.rho * u
    cs = np.sqrt(10.0 / 9.0 * u)
    alpha = np.ones(n)
    rows = rng.choice(n, 700, replace=False)
    print(f"mixed forces: periodic box, n={n}, {len(rows)} rows", flush=True)
    ref = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", u=u, c_s=cs, alpha=alpha,
                            alpha_u=0.05, cond_signal=1, rows=rows)  # fmt: skip
    got = forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, Neighbours(nb.offset, nb.index, box), "cubic", u=u,
                       c_s=cs, alpha=alpha, shock=ShockCapturing(alpha_u=0.05, conductivity_signal=1), rows=rows,
                       precision="mixed")  # fmt: skip
    _scaled("periodic rows: acc", ref.acc, got.acc, 1e-5)
    _scaled("periodic rows: dudt", ref.dudt, got.dudt, 1e-5)


def test_mixed_gradients(rng):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import gradients_numba
    from test_gpu_kernels import _moving_state

    from oracle import sph as osph

    pos, vel, mass, d, p, u, cs, alpha, nb = _moving_state(rng)
    acc = rng.normal(0, 1, pos.shape)
    nbg = Neighbours(nb.offset, nb.index, None)
    print(f"mixed gradients: n={len(pos)}", flush=True)
    for a in (acc, None):
        ref = osph.gradients_numba(pos, vel, a, mass, d.h, nb, "cubic")
        got = gradients_numba(pos, vel, a, mass, d.h, nbg, "cubic", precision="mixed")
        tag = "acc given" if a is not None else "acc zero"
        _scaled(f"div v ({tag})", ref.div_v, got.div_v, 1e-4)
        _scaled(f"|curl v| ({tag})", ref.curl_v, got.curl_v, 1e-4)
        _scaled(f"d(div v)/dt ({tag})", ref.ddivv_dt, got.ddivv_dt, 1e-4)


def test_mixed_coupled_stream_crash():
    from dataclasses import replace

    from nebula.ics import stream_crash
    from nebula.sim import coupled

    from oracle import coupled as oc

    ic = stream_crash(8192, length=200.0, seed=11)
    base = coupled.baseline_config(balsara=True)
    cfg = replace(base, cooling_powerlaw=True)
    cfg_m = replace(cfg, hydro=replace(cfg.hydro, precision="mixed"))
    print("mixed coupled step: stream crash n=8192, Balsara + power-law cooling, 4 steps", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_m)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    e0 = oc.ledger(ost, cfg)
    for k in range(4):
        dt = coupled.stable_dt(st, cfg_m)
        odt = oc.stable_dt(ost, cfg)
        check(f"step {k}: stable_dt (relative 1e-5)", odt, dt, rtol=1e-5)
        st = coupled.step(st, cfg_m, odt)
        ost = oc.step(ost, cfg, odt)
    got = st.to_numpy()
    h_ref = ost.caller_order("h")
    dx = np.linalg.norm(got["x"] - ost.caller_order("x"), axis=1) / h_ref
    ok = float(dx.max()) <= 1e-6
    print(("PASS " if ok else "FAIL ") + f"x: max |dx| / h = {float(dx.max()):.3g} (bound 1e-6)")
    assert ok
    _relative("u", ost.caller_order("u"), got["u"], 1e-4)
    for k in ("h", "rho"):
        _relative(k, ost.caller_order(k), got[k], 1e-5)
    _scaled("a_hydro", ost.caller_order("a_hydro"), got["a_hydro"], 1e-4)
    lg, lo = coupled.ledger(st, cfg_m), oc.ledger(ost, cfg)
    scale = lo["abs_h"]
    check("energy ledger (1e-9 of sum m|H|)", lo["energy"] / scale, lg["energy"] / scale, atol=1e-9)
    drift_m = (lg["energy"] - e0["energy"]) / scale  # the ledger's energy includes E_radiated
    drift_o = (lo["energy"] - e0["energy"]) / scale
    print(f"energy drift over 4 steps / sum m|H|: FP64 {drift_o:.3g}, mixed {drift_m:.3g}")
    ok = abs(drift_m) <= abs(drift_o) + 1e-9
    print(("PASS " if ok else "FAIL ") + "energy drift of the mixed run within the FP64 run's + 1e-9")
    assert ok
