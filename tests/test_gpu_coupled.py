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
    for k in fields:
        bit_equal(f"state.{k}", getattr(ost, k), got[k])
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


def test_stream_crash_cooling_balsara():
    from dataclasses import replace

    from nebula.ics import stream_crash
    from nebula.sim import coupled

    ic = stream_crash(8192, length=200.0, seed=20261014)
    base = coupled.baseline_config(balsara=True)
    cfg = replace(base, cooling_powerlaw=True)
    print("coupled: 8192-particle stream crash, power-law cooling + Balsara, 4 steps", flush=True)
    st, ost = _run(ic, cfg, 4)
    _compare(st, ost, cfg)


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
    check("switch+cond+recon: x", ost.x, got["x"], rtol=1e-12, per_vector=True)
    check("switch+cond+recon: a_hydro", ost.a_hydro, got["a_hydro"], rtol=1e-9, per_vector=True)
    check("switch+cond+recon: rho", ost.rho, got["rho"], rtol=1e-12)
    check("switch+cond+recon: alpha", ost.alpha, got["alpha"], rtol=1e-9, atol=1e-15)
