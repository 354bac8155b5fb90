"""Parity at the benchmark's full size: the 1,048,576-particle stream crash (two streams, seed
20261014) with Monaghan AV + the Balsara limiter and power-law cooling.

- FP64: initialise and one coupled step on the GPU equal the FP64 oracle's bit for bit (every field
  of the state), the ledger sums to 1e-13 of sum m|H|.
- Mixed precision: the density pass's tight rows are the reference's lists at the h it solved for
  (neighbours.py:240-281), exactly; h and rho agree with the FP64 pass on the same state within
  the mixed mode's tolerance (1e-5 relative).
- Stepper (one host synchronisation per step) against coupled.step over three steps at full size:
  bit for bit.
- Momentum: the hydro accelerations of one pass carry no net force, sum m a = 0, to 1e-12 of
  sum m |a| in FP64 and 1e-6 in mixed precision (FP32 pair terms; the neighbour mass multiplies
  them in FP32 in each row).
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

N = 1 << 20
FIELDS = ("x", "p", "u", "h", "rho", "alpha", "a_hydro", "dudt", "vsig_max", "c_s")


def _config(precision):
    from dataclasses import replace

    from nebula.sim import coupled

    return replace(coupled.baseline_config(balsara=True, precision=precision), cooling_powerlaw=True)


def _workload():
    from nebula.ics import stream_crash

    return stream_crash(N, seed=20261014)


def _net_force(st):
    m = st.mass.double()
    a = st.a_hydro
    net = (m[:, None] * a).sum(0).abs().max().item()
    scale = (m * a.norm(dim=1)).sum().item()
    return net / scale


def test_fp64_step_bit_identical_at_full_size():
    from nebula.sim import coupled

    from oracle import coupled as oc

    ic = _workload()
    cfg = _config("fp64")
    print(f"full size: FP64 initialise + 1 step, {N} particles, GPU and oracle", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    ost = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    dt, odt = coupled.stable_dt(st, cfg), oc.stable_dt(ost, cfg)
    check("stable_dt bit-identical", odt, dt)
    st = coupled.step(st, cfg, dt)
    ost = oc.step(ost, cfg, odt)
    got = st.to_numpy()
    bit_equal("state.perm", ost.perm, st.perm.cpu().numpy())
    for k in FIELDS:
        bit_equal(f"state.{k}", ost.caller_order(k), got[k])
    lg, lo = coupled.ledger(st, cfg), oc.ledger(ost, cfg)
    scale = lo["abs_h"]
    check("ledger energy (1e-13 of sum m|H|)", lo["energy"] / scale, lg["energy"] / scale, atol=1e-13)
    check("ledger e_radiated (1e-13 of sum m|H|)", lo["e_radiated"] / scale, lg["e_radiated"] / scale, atol=1e-13)
    net = _net_force(st)
    print(f"FP64 net hydro force / sum m|a| = {net:.3g}")
    check("FP64: no net hydro force (1e-12 of sum m|a|)", True, net < 1e-12)


def test_mixed_lists_exact_and_fields_at_full_size():
    import torch

    from nebula import _dev
    from nebula.sim import coupled, sph

    from oracle import neighbours as on

    ic = _workload()
    cfg = _config("mixed")
    hc = cfg.hydro
    print(f"full size: mixed density pass, {N} particles, lists against the oracle", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    n = int(st.x.shape[0])
    cloud, seg_off = coupled.one_cloud(n, st.x.device)
    counts = _dev.empty((n, 2), torch.int32)
    d, _ = sph.hydro_density_mixed_device(st.x, st.mass, st.h.clone(), hc.kernel, hc.eta, 1e-4, hc.h_max, hc.skin,
                                          st.m_uniform, cloud_d=cloud, seg_off=seg_off,
                                          row_hint=st.row_counts, row_counts=counts)  # fmt: skip
    x = st.x.cpu().numpy()
    h = d.h.cpu().numpy()
    ref = on.find_neighbours_grid(x, h)
    got = d.neighbours.to_csr()
    bit_equal("mixed tight rows = reference lists at the solved h: offsets", ref.offset, got.offset)
    bit_equal("mixed tight rows = reference lists at the solved h: indices", ref.index, got.index)
    # the FP64 pass on the same state
    d64, _ = sph.hydro_density_device(st.x, st.mass, st.h.clone(), hc.kernel, hc.eta, 1e-4, hc.h_max, hc.skin,
                                      st.m_uniform, "fp64")  # fmt: skip
    rel_h = float(((d.h - d64.h).abs() / d64.h).max())
    rel_rho = float(((d.rho - d64.rho).abs() / d64.rho).max())
    print(f"mixed vs FP64 pass on the same state: h {rel_h:.3g}, rho {rel_rho:.3g} (max relative)")
    check("mixed h within 1e-5 of FP64", True, rel_h < 1e-5)
    check("mixed rho within 1e-5 of FP64", True, rel_rho < 1e-5)
    net = _net_force(st)
    print(f"mixed net hydro force / sum m|a| = {net:.3g}")
    check("mixed: no net hydro force (1e-6 of sum m|a|)", True, net < 1e-6)


def test_stepper_equals_coupled_step_at_full_size():
    from nebula.sim import coupled
    from nebula.sim.stepper import Stepper

    ic = _workload()
    cfg = _config("mixed")
    print(f"full size: Stepper vs coupled.step, {N} particles, 3 steps", flush=True)
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    sp = Stepper(coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg), cfg)
    for _ in range(3):
        dt = coupled.stable_dt(st, cfg)
        st = coupled.step(st, cfg, dt)
        check("dt of the step", dt, sp.step())
    for k in FIELDS + ("perm", "row_counts"):
        bit_equal(f"{k}", getattr(st, k).cpu().numpy(), getattr(sp.state, k).cpu().numpy())
    print("stepper fallbacks:", sp.fallbacks, sp.fallback_reasons)
