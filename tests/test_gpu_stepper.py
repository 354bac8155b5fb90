"""GPU tests of the single-synchronisation step driver (nebula.sim.stepper.Stepper).

- Its steps give coupled.step's bits in the mixed-precision mode: every field, the time steps
  (coupled.stable_dt of each state), the list totals, with the benchmark's physics (stream crash,
  Balsara limiter, power-law cooling, GN flow) and the learned start learning online.
- A step whose fixed capacities are too small reports it, is discarded and done again by
  coupled.step; the results are still coupled.step's.
- PipelinedStepper (step k + 1 launched before step k is read back, dt taken on the device): the
  same bits, with and without the learned start (whose network steps wait for the read-back) and
  with the gate switching the network on and off; a failed step discards the step launched after
  it and is done again by coupled.step; the launches overlap where the network does not run.
- RecordedStepper (the same step recorded once as a hipGraph, dt written to the device before
  each replay): the same bits over ten replays, with and without the learned start; the graph
  holds no memset node; an undersized recording falls back, records again and still gives
  coupled.step's bits.
"""

import copy

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FIELDS = ("x", "p", "u", "h", "rho", "alpha", "a_hydro", "dudt", "vsig_max", "c_s", "omega", "perm", "row_counts")


def _config(learned_model=None):
    from dataclasses import replace

    from nebula.sim import coupled

    cfg = replace(coupled.baseline_config(balsara=True, precision="mixed"), cooling_powerlaw=True)
    if learned_model is not None:
        cfg = replace(cfg, hydro=replace(cfg.hydro, learned_h=True, learned_model=learned_model))
    return cfg


def _compare(label, a, b):
    for f in FIELDS:
        bit_equal(f"{label}: {f}", getattr(a, f).cpu().numpy(), getattr(b, f).cpu().numpy())
    check(f"{label}: t", a.t, b.t)
    check(f"{label}: radiated energy", float(a.e_radiated), float(b.e_radiated))
    check(f"{label}: search entries", a.n_search_entries, b.n_search_entries)
    check(f"{label}: tight entries", a.n_tight_entries, b.n_tight_entries)


def _trained_model(ic, cfg, gated=False):
    from dataclasses import replace

    from nebula.sim import coupled
    from nebula.sim.learned_h import LearnedH, PairLog, StartGate

    model = LearnedH(seed=5)
    model.collect = PairLog(stride=2)
    c = replace(cfg, hydro=replace(cfg.hydro, learned_model=model))
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c)
    for _ in range(3):
        st = coupled.step(st, c, coupled.stable_dt(st, c))
    f, t = model.collect.tensors()
    model.collect = None
    model.fit_normalisation(f, t)
    model.train(f, t, steps=60, batch=2048, record=False)
    model.start_online(stride=2, every=3, steps=2, batch=1024)
    if gated:  # probes every 4 steps, so that 10 steps see the gate switch both ways
        model.gate = StartGate(on_above=-1.0, off_below=0.0, probe_every=4)
    return model


@pytest.mark.parametrize("driver", ["stepper", "pipelined"])
@pytest.mark.parametrize("learned", [False, True, "gated"])
def test_stepper_equals_coupled_step(learned, driver):
    """With a gated learned start (probes every 4 steps, on whenever a probe finds a slow row) the
    gate's decisions are part of what must match: the stepper redoes a failed step with the
    decision it had made, and observes it once."""
    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import PipelinedStepper, Stepper

    ic = stream_crash(16384, length=200.0, seed=20261014)
    base = _config()
    model_a = _trained_model(ic, base, gated=learned == "gated") if learned else None
    model_b = copy.deepcopy(model_a) if learned else None
    cfg_a, cfg_b = _config(model_a), _config(model_b)
    print(f"{driver}: 16384-particle stream crash, learned start {learned}, 10 steps", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_a)
    cls = PipelinedStepper if driver == "pipelined" else Stepper
    sp = cls(coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_b), cfg_b)
    dts, sdts = [], []
    for _ in range(10):
        dt = coupled.stable_dt(st, cfg_a)
        dts.append(dt)
        st = coupled.step(st, cfg_a, dt)
        sdts.append(sp.step())
    bit_equal("time steps", np.array(dts), np.array(sdts))
    _compare("stepper vs coupled.step", st, sp.state)
    # the first steps from the initial conditions may outgrow their rows or search (the eager
    # step then regathers or rebuilds): those steps are done again by coupled.step
    print("fallbacks:", sp.fallbacks, sp.fallback_reasons)
    check("most steps without a fallback", True, sp.fallbacks <= 3)
    if driver == "pipelined":
        print("overlapped launches:", sp.overlapped, "discarded:", sp.discarded)
        if learned is False:
            check("most launches overlapped their predecessor", True, sp.overlapped >= 5)
        elif learned == "gated":  # the probes (plain starts) overlap; network steps wait
            check("the plain-start launches overlapped", True, sp.overlapped >= 1)
    if learned:
        bit_equal("online-trained weights", model_a.params.cpu().numpy(), model_b.params.cpu().numpy())
    if learned == "gated":
        print("gate:", model_a.gate.stats(), model_b.gate.stats())
        check("the gates made the same decisions", model_a.gate.stats(), model_b.gate.stats())
        check("the gate ran the network on some steps and probed on others", True,
              model_a.gate.steps_network > 0 and model_a.gate.probes >= 2)  # fmt: skip


@pytest.mark.parametrize("driver", ["stepper", "pipelined"])
def test_stepper_overflow_falls_back_to_coupled_step(driver):
    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import PipelinedStepper, Stepper

    ic = stream_crash(16384, length=200.0, seed=7)
    cfg = _config()
    print(f"{driver}: capacities a third of the need, 4 steps", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    cls = PipelinedStepper if driver == "pipelined" else Stepper
    sp = cls(coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg), cfg, headroom=0.3)
    for _ in range(4):
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        sp.step()
    print("fallbacks:", sp.fallbacks, sp.fallback_reasons)
    check("the undersized first step fell back", True, sp.fallbacks >= 1)
    check("reported as a capacity overflow", True,
          bool(set(sp.fallback_reasons[0]) & {"gather rows", "extras", "tight index", "grid cells"}))
    _compare("stepper (with fallback) vs coupled.step", st, sp.state)


@pytest.mark.parametrize("learned", [False, True])
def test_recorded_stepper_equals_coupled_step(learned):
    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import RecordedStepper

    ic = stream_crash(16384, length=200.0, seed=20261014)
    base = _config()
    model_a = _trained_model(ic, base) if learned else None
    model_b = copy.deepcopy(model_a) if learned else None
    cfg_a, cfg_b = _config(model_a), _config(model_b)
    print(f"recorded stepper: 16384-particle stream crash, learned start {learned}, 10 replays", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_a)
    for _ in range(2):  # past the first steps, whose rows still grow from the initial conditions
        st = coupled.step(st, cfg_a, coupled.stable_dt(st, cfg_a))
    st_b = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_b)
    for _ in range(2):
        st_b = coupled.step(st_b, cfg_b, coupled.stable_dt(st_b, cfg_b))
    rs = RecordedStepper(st_b, cfg_b)
    print("recorded graph nodes:", rs.node_counts)
    check("no memset node in the recorded step", 0, rs.node_counts["memset"])
    dts, rdts = [], []
    for _ in range(10):
        dt = coupled.stable_dt(st, cfg_a)
        dts.append(dt)
        st = coupled.step(st, cfg_a, dt)
        rdts.append(rs.step())
    bit_equal("time steps", np.array(dts), np.array(rdts))
    _compare("recorded stepper vs coupled.step", st, rs.state)
    print("fallbacks:", rs.fallbacks, rs.fallback_reasons, "replays", rs.replays)
    check("replays without a fallback", True, rs.fallbacks <= 2)
    if learned:
        bit_equal("online-trained weights", model_a.params.cpu().numpy(), model_b.params.cpu().numpy())


def test_recorded_stepper_overflow_records_again():
    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import RecordedStepper

    ic = stream_crash(16384, length=200.0, seed=7)
    cfg = _config()
    print("recorded stepper: capacities a third of the need, 4 replays", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    rs = RecordedStepper(coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg), cfg, headroom=0.3)
    for _ in range(4):
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        rs.step()
    print("fallbacks:", rs.fallbacks, rs.fallback_reasons)
    check("the undersized recording fell back", True, rs.fallbacks >= 1)
    _compare("recorded stepper (with fallback) vs coupled.step", st, rs.state)


def test_recorded_stepper_records_again_when_scratch_moves():
    """Eager calls between replays that release or grow the library's scratch (nb_ctx_trim, the
    reference's ledger pattern, a step of a larger cloud) move buffers the graph points at: the
    next replay must be recorded again first, and the results stay coupled.step's."""
    from nebula import _hip
    from nebula.ics import stream_crash
    from nebula.sim import coupled
    from nebula.sim.stepper import RecordedStepper

    ic = stream_crash(16384, length=200.0, seed=20261014)
    big = stream_crash(65536, length=400.0, seed=3)
    cfg = _config()
    print("recorded stepper: ledger and a 65536-particle eager step between replays", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for _ in range(2):
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
    st_b = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    for _ in range(2):
        st_b = coupled.step(st_b, cfg, coupled.stable_dt(st_b, cfg))
    rs = RecordedStepper(st_b, cfg)
    dts, rdts = [], []
    for k in range(6):
        dt = coupled.stable_dt(st, cfg)
        dts.append(dt)
        st = coupled.step(st, cfg, dt)
        rdts.append(rs.step())
        if k == 1:
            # the arena is released first, so that the eager calls below grow it whatever ran before
            g0 = _hip.generation()
            _hip.trim()
            coupled.ledger(rs.state, cfg)
            sb = coupled.initialise(big.x, big.v, big.u, big.mass, big.h, cfg)
            coupled.step(sb, cfg, coupled.stable_dt(sb, cfg))
            check("the scratch generation moved", True, _hip.generation() != g0)
    print("re-recordings:", rs.rerecords, "fallbacks:", rs.fallbacks, rs.fallback_reasons)
    check("the replay after the eager calls was recorded again", True, rs.rerecords >= 1)
    bit_equal("time steps", np.array(dts), np.array(rdts))
    _compare("recorded stepper (scratch moved) vs coupled.step", st, rs.state)
