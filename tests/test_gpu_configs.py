"""The benchmark configurations beyond the 1M stream crash, on the GPU (mixed precision against
FP64, with the error bounds of the mixed mode):

- the 262,144-particle sphere with power-law cooling: 4 coupled steps at 32,768 particles against
  the FP64 oracle, and one hydro pass at the full 262,144 against the FP64 GPU pass (itself
  bit-identical to the oracle, tests/test_gpu_full_size.py) on the same state;
- the 64-cloud ensemble (counts log-uniform in [8192, 262144], pericentres 1000-4000, radii
  100-200): one batched step, every cloud bit-identical to its solo mixed run, and each cloud
  within the mixed bounds of its FP64 solo run;
- the torus with Shakura-Sunyaev viscosity and cooling at 4,096 particles in mixed precision
  with the learned start (trained on the run's own logged pairs, learning online), against the
  FP64 oracle with the plain start, per field. The learned start only moves where Newton starts;
  h still converges to the same tolerance (1e-4 of h), so h and rho are held to that.
"""

from dataclasses import replace

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _rel(name, ref, got, bound):
    err = np.abs(got - ref) / np.abs(ref)
    print(f"{name}: median {np.median(err):.3g}, p99 {np.quantile(err, 0.99):.3g}, max {err.max():.3g} (bound {bound})")
    check(f"{name} within {bound} relative", True, bool(np.all(err <= bound)))
    return float(err.max())


def _rms(name, ref, got, bound):
    ref = ref.reshape(len(ref), -1)
    got = got.reshape(len(got), -1)
    rms = float(np.sqrt(np.mean(np.sum(ref**2, axis=1))))
    err = np.linalg.norm(got - ref, axis=1) / rms
    print(f"{name}: median {np.median(err):.3g}, p99 {np.quantile(err, 0.99):.3g}, max {err.max():.3g} of rms "
          f"(bound {bound})")  # fmt: skip
    check(f"{name} within {bound} of rms", True, bool(np.all(err <= bound)))
    return float(err.max())


def _x_over_h(ref_x, got_x, h, bound=1e-6):
    d = np.linalg.norm(got_x - ref_x, axis=1) / h
    print(f"x: max |dx| / h {d.max():.3g} (bound {bound})")
    check(f"x within {bound} of h", True, bool(d.max() <= bound))


def test_sphere_cooling_mixed_vs_fp64_oracle():
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    from oracle import coupled as oc

    ic = jittered_sphere_on_parabola(32768, r0=2100.0)
    c64 = replace(coupled.baseline_config(), cooling_powerlaw=True)
    cm = replace(coupled.baseline_config(precision="mixed"), cooling_powerlaw=True)
    print("sphere: 32768 particles from r = 2100 (pericentre 2000), power-law cooling, mixed vs FP64 oracle, 4 steps",
          flush=True)  # fmt: skip
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cm)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c64)
    for k in range(4):
        odt = oc.stable_dt(ost, c64)
        check(f"step {k}: stable_dt (1e-5 relative)", odt, coupled.stable_dt(st, cm), rtol=1e-5)
        st = coupled.step(st, cm, odt)
        ost = oc.step(ost, c64, odt)
    got = st.to_numpy()
    _x_over_h(ost.caller_order("x"), got["x"], ost.caller_order("h"))
    _rel("u", ost.caller_order("u"), got["u"], 1e-4)
    _rel("h", ost.caller_order("h"), got["h"], 1e-5)
    _rel("rho", ost.caller_order("rho"), got["rho"], 1e-5)
    _rms("a_hydro", ost.caller_order("a_hydro"), got["a_hydro"], 1e-4)
    _rms("du/dt", ost.caller_order("dudt"), got["dudt"], 1e-4)
    check("energy was radiated by the cooling", True, ost.e_radiated > 0.0)
    check("radiated energy (1e-5 relative)", ost.e_radiated, float(st.e_radiated), rtol=1e-5)


def test_sphere_full_size_one_pass_mixed_vs_fp64():
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled

    ic = jittered_sphere_on_parabola(262144, r0=2100.0)
    c64 = replace(coupled.baseline_config(), cooling_powerlaw=True)
    cm = replace(coupled.baseline_config(precision="mixed"), cooling_powerlaw=True)
    print("sphere: 262144 particles, one hydro pass, mixed vs FP64 on the same state", flush=True)
    a = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c64).to_numpy()
    b = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cm).to_numpy()
    _rel("h", a["h"], b["h"], 1e-5)
    _rel("rho", a["rho"], b["rho"], 1e-5)
    _rms("a_hydro", a["a_hydro"], b["a_hydro"], 1e-4)
    _rms("du/dt", a["dudt"], b["dudt"], 1e-4)
    _rel("v_sig,max", a["vsig_max"], b["vsig_max"], 1e-6)


def test_ensemble_64_clouds_one_step():
    import torch

    from nebula.sim import coupled
    from nebula.sim.ensemble import Ensemble, ensemble_ics

    cm = coupled.baseline_config(precision="mixed")
    c64 = coupled.baseline_config()
    ics = ensemble_ics(64, 8192, 262144)
    sizes = [len(ic.x) for ic in ics]
    print(f"ensemble: 64 clouds, {sum(sizes)} particles ({min(sizes)}..{max(sizes)}), one step", flush=True)
    ens = Ensemble(ics, cm)
    dts = ens.step()
    got = ens.state.to_numpy()
    off = np.concatenate([[0], np.cumsum(sizes)])
    worst = {"h": 0.0, "rho": 0.0, "a": 0.0}
    for k, ic in enumerate(ics):
        solo = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cm, uniform_mass=False)
        check(f"cloud {k}: dt = the solo run's stable_dt", coupled.stable_dt(solo, cm), float(dts[k]))
        solo = coupled.step(solo, cm, float(dts[k])).to_numpy()
        sl = slice(off[k], off[k + 1])
        for f in ("x", "p", "u", "h", "rho", "a_hydro", "dudt"):
            bit_equal(f"cloud {k} ({sizes[k]}): {f} = solo run", solo[f], got[f][sl])
        s64 = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c64)
        s64 = coupled.step(s64, c64, float(dts[k])).to_numpy()
        worst["h"] = max(worst["h"], _rel(f"cloud {k}: h vs FP64", s64["h"], solo["h"], 1e-5))
        worst["rho"] = max(worst["rho"], _rel(f"cloud {k}: rho vs FP64", s64["rho"], solo["rho"], 1e-5))
        worst["a"] = max(worst["a"], _rms(f"cloud {k}: a_hydro vs FP64", s64["a_hydro"], solo["a_hydro"], 1e-4))
        torch.cuda.synchronize()
    print(f"ensemble of 64: largest mixed-vs-FP64 errors over the clouds {worst}")


def test_torus_mixed_learned_vs_fp64_oracle():
    from nebula.ics import torus
    from nebula.sim import coupled
    from nebula.sim.learned_h import LearnedH, PairLog

    from oracle import coupled as oc

    ic = torus(4096, seed=5)
    c64 = replace(coupled.baseline_config(ns_alpha=0.1), cooling_powerlaw=True)
    cm = replace(coupled.baseline_config(ns_alpha=0.1, precision="mixed"), cooling_powerlaw=True)
    # the network: pairs logged by 3 plain mixed steps of the same torus, then online learning
    model = LearnedH(seed=2)
    model.collect = PairLog()
    clog = replace(cm, hydro=replace(cm.hydro, learned_model=model))
    s = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, clog)
    for _ in range(3):
        s = coupled.step(s, clog, coupled.stable_dt(s, clog))
    f, t = model.collect.tensors()
    model.collect = None
    model.fit_normalisation(f, t)
    model.train(f, t, steps=100, batch=2048, record=False)
    model.start_online(stride=1, every=2, steps=4, batch=1024)
    cl = replace(cm, hydro=replace(cm.hydro, learned_h=True, learned_model=model))
    print("torus: 4096 particles, Shakura-Sunyaev + cooling, mixed + learned start vs FP64 oracle, 4 steps",
          flush=True)  # fmt: skip
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cl)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c64)
    iters = []
    for k in range(4):
        odt = oc.stable_dt(ost, c64)
        check(f"step {k}: stable_dt (1e-4 relative)", odt, coupled.stable_dt(st, cl), rtol=1e-4)
        st = coupled.step(st, cl, odt)
        ost = oc.step(ost, c64, odt)
        iters.append((float(st.iterations.double().mean().item()), float(np.mean(ost.iterations))))
    print("Newton iterations per particle (mixed + learned, FP64 plain):", iters)
    got = st.to_numpy()
    _x_over_h(ost.caller_order("x"), got["x"], ost.caller_order("h"))
    _rel("u", ost.caller_order("u"), got["u"], 1e-4)
    _rel("h (learned start: Newton tolerance 1e-4)", ost.caller_order("h"), got["h"], 1e-4)
    _rel("rho", ost.caller_order("rho"), got["rho"], 3e-4)
    _rms("a_hydro", ost.caller_order("a_hydro"), got["a_hydro"], 1e-3)
    _rms("du/dt", ost.caller_order("dudt"), got["dudt"], 1e-3)
    fb = model.last_fallbacks
    check("learned start: fallback count is reported", True, fb is not None)
