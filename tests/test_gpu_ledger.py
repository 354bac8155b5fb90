"""The reference's hot-blob ledger test on the GPU step (tests/sim/test_coupled.py:45-77).

A warm ~500-particle blob follows the r_a = 500 -> r_p = 25 orbit from apocentre through the
pericentre squeeze to t = 18,000 with the reference's hydro defaults (Wendland C4, Cullen-Dehnen
switch, velocity reconstruction, velocity-signal conductivity). The reference's bounds hold for
the GPU step in both modes:

- FP64: max |dE| / sum m|H| < 1e-5 at (eta_orb, C) = (0.00125, 0.15), falling 3-5x from
  (0.0025, 0.3); |dL|/|L| < 1e-10;
- mixed precision (FP32 pair sums): the same energy bound (the FP64 oracle's tolerance), and
  |dL|/|L| < 1e-10 also with 1 % mass noise, where each FP32 pair force must carry both masses
  symmetrically for the momentum exchange to cancel.
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import T_END_BLOB, U_BLOB, check, hot_blob_config, hot_blob_ic

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _run(eta_orb, courant, precision="fp64", mass_noise=0.0):
    from nebula.sim import coupled

    cfg = hot_blob_config(eta_orb, courant, precision=precision)
    st = coupled.initialise(*hot_blob_ic(mass_noise=mass_noise), cfg)
    led0 = coupled.ledger(st, cfg)
    out = {"e": 0.0, "l": 0.0, "u_max": 0.0, "steps": 0}
    while st.t < T_END_BLOB * (1 - 1e-14):
        st = coupled.step(st, cfg, min(coupled.stable_dt(st, cfg), T_END_BLOB - st.t))
        led = coupled.ledger(st, cfg)
        out["e"] = max(out["e"], abs(led["energy"] - led0["energy"]) / led0["abs_h"])
        out["l"] = max(out["l"], float(np.linalg.norm(led["angmom"] - led0["angmom"]) / np.linalg.norm(led0["angmom"])))
        out["steps"] += 1
        if out["steps"] % 64 == 0:
            out["u_max"] = max(out["u_max"], float(st.u.max().item()))
    out["u_max"] = max(out["u_max"], float(st.u.max().item()))
    print(f"hot blob ({precision}, eta_orb {eta_orb}, C {courant}, mass noise {mass_noise}): {out['steps']} steps, "
          f"max |dE|/sum m|H| {out['e']:.3e}, max |dL|/|L| {out['l']:.3e}, u_max/u0 {out['u_max'] / U_BLOB:.1f}",
          flush=True)  # fmt: skip
    return out


def test_hot_blob_ledger_fp64():
    fine = _run(0.00125, 0.15)
    coarse = _run(0.0025, 0.3)
    check("FP64 hot blob: max |dE| / sum m|H| < 1e-5", 0.0, fine["e"], atol=1e-5)
    check("FP64 hot blob: error ratio coarse / fine (second order)", 4.0, coarse["e"] / fine["e"], atol=1.0)
    check("FP64 hot blob: max |dL| / |L| < 1e-10", 0.0, fine["l"], atol=1e-10)
    check("FP64 hot blob: heated at pericentre (u rose > 10x)", True, fine["u_max"] > 10 * U_BLOB)


def test_hot_blob_ledger_mixed():
    fine = _run(0.00125, 0.15, precision="mixed")
    check("mixed hot blob: max |dE| / sum m|H| < 1e-5", 0.0, fine["e"], atol=1e-5)
    check("mixed hot blob: max |dL| / |L| < 1e-10", 0.0, fine["l"], atol=1e-10)
    noisy = _run(0.00125, 0.15, precision="mixed", mass_noise=0.01)
    check("mixed hot blob, 1% mass noise: max |dE| / sum m|H| < 1e-5", 0.0, noisy["e"], atol=1e-5)
    check("mixed hot blob, 1% mass noise: max |dL| / |L| < 1e-10", 0.0, noisy["l"], atol=1e-10)


def test_stream_crash_angular_momentum_with_mass_noise():
    """The benchmark's stream crash (1 % mass noise, Balsara limiter, power-law cooling) at
    16,384 particles, 1000 steps through the single-synchronisation stepper: the mixed step's
    angular-momentum ledger stays within the reference's 1e-10 (tests/sim/test_coupled.py:56),
    as the FP64 step's does."""
    from dataclasses import replace

    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import Stepper

    ic = stream_crash(16384, length=200.0, seed=20261014)
    check("the workload's masses differ", True, float(np.std(ic.mass) / np.mean(ic.mass)) > 0.005)
    out = {}
    for precision in ("mixed", "fp64"):
        cfg = replace(coupled.baseline_config(balsara=True, precision=precision), cooling_powerlaw=True)
        st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
        l0 = coupled.ledger(st, cfg)
        sp = Stepper(st, cfg) if precision == "mixed" else None
        e_max = l_max = 0.0
        for k in range(1000):
            if sp is not None:
                sp.step()
                st = sp.state
            else:
                st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
            if k % 10 == 9:
                lg = coupled.ledger(st, cfg)
                e_max = max(e_max, abs(lg["energy"] - l0["energy"]) / l0["abs_h"])
                l_max = max(l_max, float(np.linalg.norm(lg["angmom"] - l0["angmom"]) / np.linalg.norm(l0["angmom"])))
        out[precision] = (e_max, l_max)
        print(f"stream crash 16384, {precision}, 1000 steps: max |dE|/sum m|H| {e_max:.3e}, max |dL|/|L| {l_max:.3e}",
              flush=True)  # fmt: skip
    check("mixed: max |dL| / |L| < 1e-10", 0.0, out["mixed"][1], atol=1e-10)
    check("FP64: max |dL| / |L| < 1e-10", 0.0, out["fp64"][1], atol=1e-10)
    check("mixed energy error within the FP64 run's (x 1.01 + 1e-8)", True,
          out["mixed"][0] <= 1.01 * out["fp64"][0] + 1e-8)  # fmt: skip
