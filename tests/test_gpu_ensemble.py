"""GPU tests of ensembles of clouds (nebula.sim.ensemble) and of the launch-shape autotuner
(nebula.sim.autotune).

- An ensemble of uneven clouds, two of them overlapping in space, stepped together with a time
  step per cloud, reproduces solo runs of each cloud bit for bit: the same time steps, and the
  same x, p, u, h, rho, accelerations and du/dt. The segmented grid gives each cloud the grid a
  solo run would build, so even the order of every FP32 pair sum is the solo run's.
- The workgroup size of the mixed-precision kernels changes no result bit.
- The autotuner returns one of its candidates, and a step with the tuned grid stays within the
  mixed mode's density tolerance of the default grid's.
"""

from dataclasses import replace

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FIELDS = ("x", "p", "u", "h", "rho", "a_hydro", "dudt")


def _clouds():
    from nebula.ics import jittered_sphere_on_parabola

    a = jittered_sphere_on_parabola(3000, radius=150.0, r_p=2000.0, seed=11)
    b = jittered_sphere_on_parabola(5200, radius=180.0, r_p=3000.0, seed=12)
    # c sits inside a's volume (same orbit, offset by a third of a's radius): the clouds overlap
    c = jittered_sphere_on_parabola(1700, radius=120.0, r_p=2000.0, seed=13)
    c = replace(c, x=c.x + np.array([50.0, 0.0, 0.0]))
    return [a, b, c]


def _solo(ic, cfg, dts):
    from nebula.sim import coupled

    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg, uniform_mass=False)
    got = []
    for dt in dts:
        d = coupled.stable_dt(st, cfg)
        got.append(d)
        st = coupled.step(st, cfg, dt if dt is not None else d)
    return st.to_numpy(), got


def test_ensemble_reproduces_solo_runs():
    from nebula.sim import coupled
    from nebula.sim.ensemble import Ensemble

    cfg = coupled.baseline_config(precision="mixed")
    ics = _clouds()
    steps = 4
    print(f"ensemble: clouds of {[len(ic.x) for ic in ics]} particles, {steps} steps", flush=True)
    ens = Ensemble(ics, cfg)
    dts = [ens.step() for _ in range(steps)]
    for k, ic in enumerate(ics):
        print(f"solo run of cloud {k}", flush=True)
        solo, solo_dt = _solo(ic, cfg, [None] * steps)
        bit_equal(f"cloud {k}: time steps", np.array(solo_dt), np.array([d[k] for d in dts]))
        mine = ens.cloud(k)
        for f in FIELDS:
            bit_equal(f"cloud {k}: {f}", solo[f], mine[f])
    check("clouds advanced by their own steps", True, len(set(np.round(ens.t, 12))) == len(ics))


def test_tile_size_changes_no_bit():
    from nebula import _hip
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    cfg = coupled.baseline_config(precision="mixed")
    ic = jittered_sphere_on_parabola(6000, seed=21)
    out = {}
    try:
        for tile in (256, 128, 64):
            print(f"mixed step with {tile}-thread workgroups", flush=True)
            _hip.set_tile(tile)
            st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
            st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
            out[tile] = st.to_numpy()
    finally:
        _hip.set_tile(256)
    for tile in (128, 64):
        for f in FIELDS:
            bit_equal(f"tile {tile} vs 256: {f}", out[256][f], out[tile][f])


def test_autotune_choice_and_tolerance():
    from nebula import _hip
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import autotune, coupled

    cfg = coupled.baseline_config(precision="mixed")
    ic = jittered_sphere_on_parabola(20000, seed=31)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    try:
        print("autotune: cell factors (1.0, 1.5) x tiles (64, 256)", flush=True)
        ch = autotune.tune(st, cfg, factors=(1.0, 1.5), tiles=(64, 256), repeats=1)
        print("autotune table (ms per hydro pass):", {k: round(v, 3) for k, v in ch.table.items()})
        check("choice is a candidate", True, (ch.cell_factor, ch.tile) in ch.table)
        check("choice is the fastest", ch.ms, min(ch.table.values()))
        dt = coupled.stable_dt(st, cfg)
        ref = coupled.step(st, cfg, dt).to_numpy()
        tuned = coupled.step(autotune.apply(st, autotune.Choice(1.5, 64, 0.0, {})), cfg, dt).to_numpy()
    finally:
        _hip.set_tile(256)
    check("h with a 1.5x grid cell vs the default grid", ref["h"], tuned["h"], rtol=1e-5)
    check("rho with a 1.5x grid cell vs the default grid", ref["rho"], tuned["rho"], rtol=1e-5)
