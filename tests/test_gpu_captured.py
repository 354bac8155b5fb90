"""GPU tests of the ensemble step recorded as a hipGraph (nebula.sim.captured).

- Replays of the recorded step give the eager ensemble step's bits: every field of every cloud
  and every cloud's dt, over several replays with unrelated GPU work and host reads in between
  (a graph node that misbehaves on replays after the first would show here).
- The recorded graph holds kernel and copy nodes only (no memset node).
- A replay whose fixed capacities are too small reports it, restores its starting state, redoes
  the step eagerly and records again: the results are still the eager ones.
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FIELDS = ("x", "p", "u", "h", "rho", "a_hydro", "dudt", "alpha", "perm")


def _clouds():
    from dataclasses import replace

    from nebula.ics import jittered_sphere_on_parabola

    a = jittered_sphere_on_parabola(3000, radius=150.0, r_p=2000.0, seed=11)
    b = jittered_sphere_on_parabola(5200, radius=180.0, r_p=3000.0, seed=12)
    c = jittered_sphere_on_parabola(1700, radius=120.0, r_p=2000.0, seed=13)
    c = replace(c, x=c.x + np.array([50.0, 0.0, 0.0]))
    return [a, b, c]


def _compare(k, eager, cap):
    import torch

    for f in FIELDS:
        bit_equal(f"replay {k}: {f}", getattr(eager.state, f).cpu().numpy(), cap.S[f].cpu().numpy())
    torch.cuda.synchronize()


def test_replayed_steps_equal_eager_steps():
    import torch

    from nebula.sim import coupled
    from nebula.sim.captured import CapturedEnsemble
    from nebula.sim.ensemble import Ensemble

    cfg = coupled.baseline_config(precision="mixed")
    print("captured ensemble: 3 clouds, one eager step each, then record", flush=True)
    eager = Ensemble(_clouds(), cfg)
    eager.step()
    other = Ensemble(_clouds(), cfg)
    other.step()
    cap = CapturedEnsemble(other)
    nodes = cap.node_counts()
    print("recorded graph nodes:", nodes)
    check("no memset node in the recorded step", 0, nodes.get("memset", 0))
    check("kernel nodes recorded", True, nodes.get("kernel", 0) > 20)
    for k in range(4):
        d_eager = eager.step()
        d_cap = cap.step()
        bit_equal(f"replay {k}: dt of each cloud", d_eager, d_cap)
        _compare(k, eager, cap)
        # unrelated device work and a host read between replays
        z = torch.randn(1 << 16, device="cuda")
        float((z * z).sum().item())
    print("fallbacks:", cap.fallbacks, cap.fallback_reasons)
    check("no replay fell back to an eager step", 0, cap.fallbacks)
    bit_equal("clocks", eager.t, other.t)


def test_overflow_falls_back_to_an_eager_step():
    from nebula.sim import coupled
    from nebula.sim.captured import CapturedEnsemble
    from nebula.sim.ensemble import Ensemble

    cfg = coupled.baseline_config(precision="mixed")
    print("captured ensemble with a tight-index capacity far too small", flush=True)
    eager = Ensemble(_clouds(), cfg)
    eager.step()
    other = Ensemble(_clouds(), cfg)
    other.step()
    cap = CapturedEnsemble(other, headroom=0.25)  # a quarter of the sizes the eager step needed
    d_eager = eager.step()
    d_cap = cap.step()
    check("the replay reported the overflow and fell back", 1, cap.fallbacks)
    bit_equal("dt after the fallback", d_eager, d_cap)
    _compare("after fallback", eager, cap)
    eager.step()
    cap.step()
    check("re-recorded with the eager step's sizes: no further fallback", 1, cap.fallbacks)
    _compare("next", eager, cap)


def test_replay_records_again_when_scratch_moves():
    """Releasing the library's scratch and an eager step of a larger ensemble between replays move
    buffers the recorded graph points at: the next replay is recorded again first, with the eager
    bits."""
    from nebula import _hip
    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled
    from nebula.sim.captured import CapturedEnsemble
    from nebula.sim.ensemble import Ensemble

    cfg = coupled.baseline_config(precision="mixed")
    print("captured ensemble: a 40000-particle eager ensemble step between replays", flush=True)
    eager = Ensemble(_clouds(), cfg)
    eager.step()
    other = Ensemble(_clouds(), cfg)
    other.step()
    cap = CapturedEnsemble(other)
    for k in range(3):
        bit_equal(f"replay {k}: dt of each cloud", eager.step(), cap.step())
        _compare(k, eager, cap)
        if k == 0:
            # the arena is released first, so that the eager step below grows it whatever ran before
            g0 = _hip.generation()
            _hip.trim()
            Ensemble([jittered_sphere_on_parabola(40000, radius=200.0, r_p=2500.0, seed=21)], cfg).step()
            check("the scratch generation moved", True, _hip.generation() != g0)
    check("recorded again after the scratch moved", True, cap.rerecords >= 1)
    check("no replay fell back to an eager step", 0, cap.fallbacks)
