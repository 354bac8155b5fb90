"""Throughput of the SPH step on one MI355X: particle-steps/s of coupled.step at 1M particles.

Workload (the 1M-particle configuration of the metric): two seeded thin cylindrical streams of
radius 20 and length 2000 (0.05 M_sun each, 1% Gaussian mass noise, 524,288 particles each)
meeting at r = 1664 M with a 160 degree opening angle at 0.056 c each, T = 1e4 K, cubic spline
with eta = 1.2, Monaghan viscosity alpha = 1, beta = 2 with a Balsara limiter, power-law cooling
on. A step is one full coupled.step (GN flow, cooling, neighbour lists, density / h solve,
equation of state, gradients for the limiter, forces, kicks) plus its stable_dt; the state is
resident in HBM before the timed region starts. value = N x steps / wall time.

Prints one JSON line. ``roofline`` covers the dominant pair kernel, timed with HIP events on
the stream the kernels run on (torch's current stream); ``cpu_baseline`` times the FP64 CPU
oracle (OpenMP, this host's cores) on the same workload.
"""

from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time
from dataclasses import replace
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (str(REPO), str(REPO / "blackhole-nebula-simulator_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK = {"fp64": 78.6, "fp32": 157.3}  # TFLOP/s, MI355X vector (dense) published peaks
HBM_PEAK = 8000.0  # GB/s

# Algorithmic FLOPs, counted from the reference's source operations (FMA = 2, div = sqrt = 1):
FLOPS_FORCE_PAIR = 83  # force pair with Monaghan viscosity, no conductivity (sph.py:542-645)
FLOPS_GRAD_PAIR = 45  # gradient pair (sph.py:848-865)
FLOPS_SEARCH_ENTRY = 10  # distance pass per search-list entry (sph.py:231-248)
FLOPS_NEWTON_ENTRY = 20  # one kernel evaluation in the Newton sums (sph.py:215-228)
FLOPS_GLF_ITER = 35  # one GLF fixed-point iteration (integrator.py:106-176)
FLOPS_PER_PARTICLE = 40  # kicks, equation of state, cooling
# VALU issue ceiling: 256 CUs x 4 SIMDs, a wave64 vector instruction every 4 cycles of a 16-lane SIMD,
# at the 2.4 GHz peak engine clock (wavefront instructions per second)
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_048_576)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-compare", action="store_true", help="skip timing the other precision alongside")
    ap.add_argument("--no-learned", action="store_true", help="mixed mode without the learned starting h")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--run-steps", type=int, default=300,
                    help="steps of the whole-run mean from the initial state (the workload's 300 steps); 0: skip")
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="steps of the energy-ledger drift runs (both precisions, from the initial state); 0: skip")
    ap.add_argument("--long-run", type=int, default=0,
                    help="instead of the headline line: step cost and device memory in windows of 50 steps over "
                         "this many steps from the initial state (the --precision mode, learned start as usual)")
    ap.add_argument("--report", choices=["sphere262k", "stream300", "torus4m"], default=None,
                    help="instead of the headline line: the error / Newton report of one of the other workloads")
    ap.add_argument("--ensemble", action="store_true",
                    help="the batched-ensemble workload (64 clouds) against sequential single-cloud runs")
    ap.add_argument("--ensemble-clouds", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (scales the PMC byte counters)")
    return ap.parse_args()


def workload(n):
    from nebula.ics import stream_crash
    from nebula.sim import coupled

    ic = stream_crash(n, seed=20261014)
    return ic


def config(precision):
    from nebula.sim import coupled

    return replace(coupled.baseline_config(balsara=True, precision=precision), cooling_powerlaw=True)


def cpu_baseline(ic, cfg, steps):
    from oracle import coupled as oc
    from oracle import lib as olib

    cores = len(os.sched_getaffinity(0))
    olib.set_threads(cores)
    st = oc.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    st = oc.step(st, cfg, oc.stable_dt(st, cfg))  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        st = oc.step(st, cfg, oc.stable_dt(st, cfg))
    wall = time.perf_counter() - t0
    n = len(ic.x)
    return {"value": n * steps / wall, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"the same {n}-particle stream crash, FP64 C/OpenMP restatement of the reference's numba "
                      f"kernels, 1 warm-up + {steps} timed coupled steps"}  # fmt: skip


def run_gpu(ic, cfg, steps, warmup, learned=False, log_stride=8, train_steps=300, driver="eager"):
    """Times ``steps`` coupled steps after ``warmup`` (no instrumentation in the timed region), then
    runs ``steps`` more with the library's kernel timing on (HIP events around each main kernel on
    its stream) and the entry-point timing. With ``learned``, the warm-up steps log (features,
    converged h) pairs (every ``log_stride``-th particle), the network is trained on them with the
    FP32 training kernels, and the timed steps start each Newton solve from its BF16 MFMA
    prediction. driver: "eager" (coupled.stable_dt + coupled.step), "stepper" (the same steps,
    bit for bit, through nebula.sim.stepper.Stepper: one host synchronisation per step).
    Returns (wall s, kernel times, entry-point times, state, info)."""
    import torch

    from nebula import _hip
    from nebula.sim import coupled
    from nebula.sim.learned_h import LearnedH, PairLog, StartGate

    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    info = {}
    if learned:
        # the gate runs the network only while the plain start leaves many rows to the slow solve
        model = LearnedH(seed=1, gate=StartGate())
        model.collect = PairLog(stride=log_stride)
        cfg_log = replace(cfg, hydro=replace(cfg.hydro, learned_model=model))
        plain_iters = []
        for _ in range(max(warmup, 2)):
            st = coupled.step(st, cfg_log, coupled.stable_dt(st, cfg_log))
            plain_iters.append(float(st.iterations.double().mean().item()))
        feat, target = model.collect.tensors()
        model.collect = None
        t0 = time.perf_counter()
        model.fit_normalisation(feat, target)
        losses = model.train(feat, target, steps=train_steps, batch=8192, lr=3e-3)
        torch.cuda.synchronize()
        info = {"pairs": int(target.numel()), "logged_steps": len(plain_iters), "train_steps": train_steps,
                "train_s": time.perf_counter() - t0, "loss_first": losses[0], "loss_last": losses[-1],
                "newton_iters_plain": sum(plain_iters) / len(plain_iters)}  # fmt: skip
        online = dict(stride=max(1, len(ic.x) // 16384), every=8, steps=4)
        model.start_online(**online)  # keeps learning from every step's own pairs
        info["online"] = online
        cfg = replace(cfg, hydro=replace(cfg.hydro, learned_h=True, learned_model=model))
        info["_cfg"] = cfg
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))  # first learned step: warm-up
    else:
        for _ in range(warmup):
            st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
    if driver in ("stepper", "pipelined"):
        from nebula.sim.stepper import PipelinedStepper, Stepper

        sp = (PipelinedStepper if driver == "pipelined" else Stepper)(st, cfg)
        for _ in range(2):  # warm-up of the fixed-capacity passes
            sp.step()

        def advance():
            sp.step()
            return sp.state
    else:
        holder = [st]

        def advance():
            holder[0] = coupled.step(holder[0], cfg, coupled.stable_dt(holder[0], cfg))
            return holder[0]

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        advance()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _hip.start_kernel_timing()
    _hip.start_timing()
    iters, fallbacks = [], 0
    for _ in range(steps):
        st = advance()
        iters.append(st.iterations.double().mean())
        fallbacks += st.learned_fallbacks if learned else 0
    torch.cuda.synchronize()
    entry = _hip.stop_timing()
    kern = _hip.stop_kernel_timing()
    info["newton_iters"] = float(sum(float(x) for x in iters) / len(iters))
    info["driver"] = driver
    if driver in ("stepper", "pipelined"):
        info["stepper_fallbacks"] = sp.fallbacks
    if driver == "pipelined":
        info["overlapped"], info["discarded"] = sp.overlapped, sp.discarded
    if learned:
        info["fallback_rate"] = fallbacks / (steps * len(ic.x))
        info["gate"] = info["_cfg"].hydro.learned_model.gate.stats()
    return wall, kern, entry, st, info


def _small_ensemble(cfg, args):
    """64 small clouds (1024..4096 particles): where launches and host work dominate a step,
    eager against recorded steps."""
    import numpy as np
    import torch

    from nebula.sim.captured import CapturedEnsemble
    from nebula.sim.ensemble import Ensemble, ensemble_ics

    ics = ensemble_ics(64, 1024, 4096, seed=7)
    total = int(sum(len(ic.x) for ic in ics))
    out = {"workload": f"64 clouds, {total} particles (1024..4096 per cloud)"}
    for mode in ("eager", "captured"):
        ens = Ensemble(ics, cfg)
        ens.step()
        runner = ens if mode == "eager" else CapturedEnsemble(ens)
        runner.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runner.step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[mode] = {"value": total * args.steps / wall, "ms_per_step": wall / args.steps * 1e3}
        del runner, ens
    out["speedup_captured"] = out["captured"]["value"] / out["eager"]["value"]
    print(f"bench: small ensemble eager {out['eager']['ms_per_step']:.2f} ms, captured "
          f"{out['captured']['ms_per_step']:.2f} ms per step", file=sys.stderr, flush=True)  # fmt: skip
    return out


def run_ensemble(args):
    """64 clouds of the plain cloud physics (particle counts log-uniform in [8192, 262144],
    pericentres 1000-4000, radii 100-200), stepped together with a step per cloud, against the same
    clouds stepped one after the other. Cell factor and tile come from the autotuner."""
    import numpy as np
    import torch

    from nebula.sim import autotune, coupled
    from nebula.sim.ensemble import Ensemble, ensemble_ics

    cfg = coupled.baseline_config(precision="mixed")
    ics = ensemble_ics(args.ensemble_clouds, 8192, 262144)
    sizes = np.array([len(ic.x) for ic in ics])
    total = int(sizes.sum())
    print(f"bench: ensemble of {len(ics)} clouds, {total} particles", file=sys.stderr, flush=True)
    ens = Ensemble(ics, cfg)
    choice = autotune.tune(ens.state, cfg, repeats=2)
    ens.state = autotune.apply(ens.state, choice)
    print(f"bench: autotuned cell factor {choice.cell_factor}, tile {choice.tile}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ens.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ens.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    batched = total * args.steps / wall
    print(f"bench: batched {batched:.4g} particle-steps/s", file=sys.stderr, flush=True)
    # the same ensemble step recorded as one hipGraph and replayed
    from nebula.sim.captured import CapturedEnsemble

    cap = CapturedEnsemble(ens)
    cap.step()  # first replay: untimed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cap.step()
    torch.cuda.synchronize()
    wall_c = time.perf_counter() - t0
    captured = {"value": total * args.steps / wall_c, "ms_per_step": wall_c / args.steps * 1e3,
                "fallbacks": cap.fallbacks, "graph_nodes": cap.node_counts(), "capacities": cap.capacities}
    print(f"bench: captured {captured['value']:.4g} particle-steps/s", file=sys.stderr, flush=True)
    del cap, ens
    small = _small_ensemble(cfg, args)
    # sequential: each cloud alone (same tuned settings), timed over a few steps after one warm-up
    seq_steps = max(2, min(5, args.steps))
    per_cloud_ms = []
    for k, ic in enumerate(ics):
        st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
        st = replace(st, cell_factor=choice.cell_factor)
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(seq_steps):
            st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        torch.cuda.synchronize()
        per_cloud_ms.append((time.perf_counter() - t0) / seq_steps * 1e3)
        if k % 16 == 15:
            print(f"bench: sequential clouds {k + 1}/{len(ics)}", file=sys.stderr, flush=True)
    seq_ms = float(np.sum(per_cloud_ms))
    sequential = total / (seq_ms / 1e3)
    out = {
        "metric": "batched-ensemble particle-steps/s (64 clouds, one launch per pass) vs 64 sequential runs",
        "value": batched,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded clouds on parabolic orbits)",
        "config": {"workload": f"ensemble of {len(ics)} clouds, {total} particles "
                               f"({int(sizes.min())}..{int(sizes.max())} per cloud)",
                   "physics": "cubic, eta 1.2, Monaghan alpha 1 beta 2, no cooling", "precision": "mixed",
                   "cell_factor": choice.cell_factor, "tile": choice.tile},
        "sequential": {"value": sequential, "ms_per_ensemble_step": seq_ms, "steps_per_cloud": seq_steps,
                       "speedup_batched": batched / sequential},
        "captured": captured,
        "small_ensemble": small,
        "autotune_ms_per_hydro_pass": {f"{f}x{t}": round(v, 3) for (f, t), v in choice.table.items()},
    }
    print(json.dumps(out))


def _err_stats(ref, got, kind):
    """median, p99 and max of the per-particle error: relative ("rel"), absolute ("abs"), or per
    vector over the rms size of ref ("rms")."""
    import numpy as np

    ref = np.asarray(ref, float)
    got = np.asarray(got, float)
    d = got - ref
    if kind == "rel":
        e = np.abs(d) / np.maximum(np.abs(ref), 1e-300)
    elif kind == "abs":
        e = np.abs(d)
    else:
        e = (np.linalg.norm(d, axis=1) if d.ndim == 2 else np.abs(d))
        e = e / float(np.sqrt(np.mean(np.sum(ref.reshape(len(ref), -1) ** 2, axis=1))))
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max())}


def parity_report(st, cfg_learned):
    """Per-field errors of the mixed mode against the FP64 path (bit-identical to the FP64 oracle)
    on the same state (one hydro pass each), and of the learned start against the plain start
    (one step each from the same state)."""
    from nebula import _dev
    from nebula.sim import coupled

    v = coupled.velocity(cfg_learned, st.x, st.p)
    a = coupled._hydro(st, config("fp64"), v, 0.0, st.a_hydro)
    b = coupled._hydro(st, config("mixed"), v, 0.0, st.a_hydro)
    host = _dev.host
    out = {"mixed_vs_fp64 (same state, one hydro pass)": {
        "h": _err_stats(host(a.h), host(b.h), "rel"), "rho": _err_stats(host(a.rho), host(b.rho), "rel"),
        "omega": _err_stats(host(a.omega), host(b.omega), "abs"),
        "acc (per vector / rms|a|)": _err_stats(host(a.a_hydro), host(b.a_hydro), "rms"),
        "dudt (/ rms)": _err_stats(host(a.dudt), host(b.dudt), "rms"),
        "vsig_max": _err_stats(host(a.vsig_max), host(b.vsig_max), "rel")}}
    del a, b
    plain = replace(cfg_learned, hydro=replace(cfg_learned.hydro, learned_h=False, learned_model=None))
    dt = coupled.stable_dt(st, plain)
    p = coupled.step(st, plain, dt)
    model = cfg_learned.hydro.learned_model
    gate, model.gate = model.gate, None  # this comparison step runs the network whatever the gate says
    q = coupled.step(st, cfg_learned, dt)
    model.gate = gate
    out["learned_vs_plain_start (one step, same state)"] = {
        "h": _err_stats(host(p.h), host(q.h), "rel"), "rho": _err_stats(host(p.rho), host(q.rho), "rel"),
        "newton_iters": [float(p.iterations.double().mean()), float(q.iterations.double().mean())]}
    return out


def drift_run(ic, cfg, steps):
    """Energy and angular-momentum ledgers (coupled.py:261-270) over ``steps`` steps from the
    initial state: the maximum over the run of |E - E0| / sum m|H| and |L - L0| / |L0| (the
    reference's coupled ledger tests bound the former by 1e-5 and the latter by 1e-10)."""
    import numpy as np

    from nebula.sim import coupled

    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    l0 = coupled.ledger(st, cfg)
    e_max = l_max = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        lg = coupled.ledger(st, cfg)
        e_max = max(e_max, abs(lg["energy"] - l0["energy"]) / l0["abs_h"])
        l_max = max(l_max, float(np.linalg.norm(lg["angmom"] - l0["angmom"]) / np.linalg.norm(l0["angmom"])))
        if k % 250 == 249:
            print(f"bench: drift run {k + 1}/{steps}, max |dE| / sum m|H| {e_max:.3g}", file=sys.stderr, flush=True)
    return {"steps": steps, "t_end": st.t, "max_energy_error": e_max, "max_angmom_error": l_max,
            "e_radiated_over_abs_h": lg["e_radiated"] / l0["abs_h"], "wall_s": time.perf_counter() - t0}


def _lockstep(ic, cfg_a, cfg_b, steps, every, compare, on_sample=None):
    """Runs ``cfg_a`` (FP64) and ``cfg_b`` from the same initial state with the FP64 run's dt at
    every step; every ``every`` steps calls compare(step, state_a, state_b)."""
    from nebula.sim import coupled

    a = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_a)
    b = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg_b)
    out = [compare(0, a, b)]
    iters_b = []
    for k in range(1, steps + 1):
        dt = coupled.stable_dt(a, cfg_a)
        a = coupled.step(a, cfg_a, dt)
        b = coupled.step(b, cfg_b, dt)
        iters_b.append(float(b.iterations.double().mean().item()))
        if k % every == 0 or k == steps:
            out.append(compare(k, a, b))
            print(f"bench: step {k}/{steps}", file=sys.stderr, flush=True)
    return a, b, out, iters_b


def _same_state(st, cfg_a, cfg_b):
    """One hydro pass of each configuration on the same state: (rho, a_hydro) errors of b vs a."""
    from nebula import _dev
    from nebula.sim import coupled

    v = coupled.velocity(cfg_a, st.x, st.p)
    a = coupled._hydro(st, cfg_a, v, 0.0, st.a_hydro)
    b = coupled._hydro(st, cfg_b, v, 0.0, st.a_hydro)
    return {"rho": _err_stats(_dev.host(a.rho), _dev.host(b.rho), "rel"),
            "acc": _err_stats(_dev.host(a.a_hydro), _dev.host(b.a_hydro), "rms"),
            "dudt": _err_stats(_dev.host(a.dudt), _dev.host(b.dudt), "rms")}


def run_report(args):
    """Reports of the other workloads (printed as one JSON line):

    - sphere262k: the plain-method cloud at 262,144 particles with power-law cooling, started at
      r = 2100 on its parabolic orbit so that 500 steps span the pericentre (2000); FP64 and mixed
      in lockstep (the FP64 run's dt), density and acceleration errors of mixed against FP64
      (bit-identical to the FP64 oracle) every 100 steps, along the two trajectories and for one
      hydro pass on the FP64 state.
    - stream300: the 1M stream crash, 300 steps, FP64 and mixed + learned h in lockstep:
      shock-heated mass fraction (T > 1e5 K), energy-ledger drift, final per-field errors.
    - torus4m: the 4,194,304-particle torus with Shakura-Sunyaev viscosity (alpha 0.1, option A)
      and power-law cooling in mixed mode: Newton iterations per particle with the plain start
      and the learned start (trained on ~2e7 pairs logged from the cloud workloads, then online;
      on every step, and gated), fallback rate, h / rho of the learned start against the plain
      start (one step from the same state)."""
    import numpy as np

    from nebula import _dev
    from nebula.ics import jittered_sphere_on_parabola, torus
    from nebula.physics import units
    from nebula.sim import coupled
    from nebula.sim.learned_h import StartGate

    rep = {"report": args.report}
    if args.report == "sphere262k":
        ic = jittered_sphere_on_parabola(262144, r0=2100.0)
        c64 = replace(coupled.baseline_config(), cooling_powerlaw=True)
        cm = replace(coupled.baseline_config(precision="mixed"), cooling_powerlaw=True)

        def compare(k, a, b):
            ha, hb = a.to_numpy(), b.to_numpy()
            return {"step": k, "t": a.t, "r_centre": float(np.linalg.norm(ha["x"].mean(axis=0))),
                    "trajectory": {"rho": _err_stats(ha["rho"], hb["rho"], "rel"),
                                   "acc": _err_stats(ha["a_hydro"], hb["a_hydro"], "rms"),
                                   "x (/ h)": _err_stats(np.zeros(len(ha["h"])),
                                                         np.linalg.norm(hb["x"] - ha["x"], axis=1) / ha["h"], "abs")},
                    "same_state": _same_state(a, c64, cm)}  # fmt: skip

        a, b, rows, it = _lockstep(ic, c64, cm, 500, 100, compare)
        la, lb = coupled.ledger(a, c64), coupled.ledger(b, cm)
        rep.update({"workload": "sphere 262,144, r0 2100 -> pericentre 2000, power-law cooling, 500 steps",
                    "samples": rows, "energy_ledger_end": [la["energy"], lb["energy"]], "abs_h": la["abs_h"]})
    elif args.report == "stream300":
        ic = workload(1 << 20)
        c64 = config("fp64")
        _, _, _, _, info = run_gpu(ic, config("mixed"), 1, 1, learned=True)
        cl = info["_cfg"]
        u_hot = units.gas_u_from_temperature(1e5)
        l0 = {}

        def compare(k, a, b):
            ha, hb = a.to_numpy(), b.to_numpy()
            la, lb = coupled.ledger(a, c64), coupled.ledger(b, cl)
            if k == 0:
                l0.update(a=la, b=lb)
            m = ha["mass"] if "mass" in ha else np.asarray(_dev.host(a.mass))
            return {"step": k, "t": a.t,
                    "shock_heated_mass_fraction": [float(np.sum(m[ha["u"] > u_hot]) / np.sum(m)),
                                                   float(np.sum(m[hb["u"] > u_hot]) / np.sum(m))],
                    "energy_drift": [(la["energy"] - l0["a"]["energy"]) / l0["a"]["abs_h"],
                                     (lb["energy"] - l0["b"]["energy"]) / l0["b"]["abs_h"]],
                    "trajectory": {"rho": _err_stats(ha["rho"], hb["rho"], "rel"),
                                   "acc": _err_stats(ha["a_hydro"], hb["a_hydro"], "rms"),
                                   "u": _err_stats(ha["u"], hb["u"], "rel")}}  # fmt: skip

        a, b, rows, it = _lockstep(ic, c64, cl, 300, 100, compare)
        rep.update({"workload": "stream crash 1,048,576, 300 steps, [FP64, mixed + learned h]", "samples": rows,
                    "newton_iters_mixed_learned": float(np.mean(it)), "hot": "T > 1e5 K"})  # fmt: skip
    else:
        n = args.n if args.n != 1_048_576 else 4_194_304
        print(f"bench: torus of {n} particles", file=sys.stderr, flush=True)
        ic = torus(n)
        base = replace(coupled.baseline_config(precision="mixed", ns_alpha=0.1), cooling_powerlaw=True)
        # The density solve (sph.py:251-297) tests convergence against tol * h_start: from a
        # random start, a first Newton step that lands near h = 0 "converges" there (the FP64
        # oracle does the same for these rows), and the next step cannot grow h back within 12
        # rebuilds. Such rows (h below 1 % of their start) get a start 1.3x larger, until none.
        relaxed = 0
        for _ in range(8):
            h_got = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, base).to_numpy()["h"]
            bad = h_got < 0.01 * ic.h
            if not bad.any():
                break
            relaxed += int(bad.sum())
            ic = replace(ic, h=np.where(bad, 1.3 * ic.h, ic.h))
        rep["collapsed_starts_moved"] = relaxed
        model, rep["training"] = _pairs_from_c1_c3()
        model.start_online(stride=max(1, n // 16384), every=8, steps=4)  # keeps learning on the torus
        cl = replace(base, hydro=replace(base.hydro, learned_h=True, learned_model=model))
        gated = copy.deepcopy(model)
        gated.gate = StartGate()
        cg = replace(base, hydro=replace(base.hydro, learned_h=True, learned_model=gated))
        steps = max(args.steps, 20)
        res = {}
        for name, c in (("plain", base), ("learned", cl), ("learned, gated", cg)):
            st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c)
            its, fb = [], 0
            t0 = time.perf_counter()
            for _ in range(steps):
                st = coupled.step(st, c, coupled.stable_dt(st, c))
                its.append(float(st.iterations.double().mean().item()))
                fb += st.learned_fallbacks
            res[name] = {"newton_iters": float(np.mean(its[1:])), "fallback_rate": fb / (steps * n),
                         "ms_per_step": (time.perf_counter() - t0) / steps * 1e3}
            if name == "plain":
                keep = st
        dt = coupled.stable_dt(keep, base)
        p = coupled.step(keep, base, dt)
        q = coupled.step(keep, cl, dt)
        if gated.gate is not None:
            res["learned, gated"]["gate"] = gated.gate.stats()
        rep.update({"workload": f"torus {n}: r_in 12, r_max 25, constant l, polytrope, 2% velocity noise, "
                                f"Shakura-Sunyaev alpha 0.1 (option A), power-law cooling; {steps} steps",
                    "runs": res, "learned_vs_plain (one step, same state)": {
                        "h": _err_stats(_dev.host(p.h), _dev.host(q.h), "rel"),
                        "rho": _err_stats(_dev.host(p.rho), _dev.host(q.rho), "rel")}})  # fmt: skip
    print(json.dumps(rep))


def _pairs_from_c1_c3(target_pairs=20_000_000):
    """The learned start of the torus report, as its configuration has it: (features, converged h)
    pairs logged by the simulator from the cloud workloads (the 4,096- and 262,144-particle
    spheres on their parabolic orbit, the latter with power-law cooling, and the 1M stream crash;
    every particle of every step, mixed precision), about ``target_pairs`` in all, then the FP32
    training kernels on them. Returns (model, what was logged and how training went)."""
    import torch

    from nebula.ics import jittered_sphere_on_parabola
    from nebula.sim import coupled
    from nebula.sim.learned_h import LearnedH, PairLog

    model = LearnedH(seed=1)
    mixed = coupled.baseline_config(precision="mixed")
    # steps per workload: C1 its 200; C2 and C3 as many as the remaining pairs need
    plan = [("sphere 4096 (C1)", jittered_sphere_on_parabola(4096), mixed, 200),
            ("sphere 262144, power-law cooling (C2)", jittered_sphere_on_parabola(262144),
             replace(mixed, cooling_powerlaw=True), 20),
            ("stream crash 1048576 (C3)", workload(1 << 20), config("mixed"), None)]  # fmt: skip
    logged, info = 0, {"sources": []}
    feats, targets = [], []
    for name, ic, cfg, steps in plan:
        n = len(ic.x)
        if steps is None:
            steps = max(1, -(-(target_pairs - logged) // n))
        print(f"bench: logging pairs from {name}, {steps} steps", file=sys.stderr, flush=True)
        model.collect = PairLog(stride=1)
        c = replace(cfg, hydro=replace(cfg.hydro, learned_model=model))
        st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c)
        for _ in range(steps):
            st = coupled.step(st, c, coupled.stable_dt(st, c))
        f, t = model.collect.tensors()
        feats.append(f)
        targets.append(t)
        logged += int(t.numel())
        info["sources"].append({"workload": name, "steps": steps, "pairs": int(t.numel())})
        model.collect = None
        del st
    f, t = torch.cat(feats), torch.cat(targets)
    del feats, targets
    t0 = time.perf_counter()
    model.fit_normalisation(f, t)
    losses = model.train(f, t, steps=2000, batch=16384, lr=3e-3)
    torch.cuda.synchronize()
    info.update({"pairs": logged, "train_steps": 2000, "batch": 16384, "train_s": time.perf_counter() - t0,
                 "loss_first": losses[0], "loss_last": losses[-1]})  # fmt: skip
    return model, info


def run_long(args):
    """Step cost and memory over a long run of the headline workload: windows of 50 steps."""
    import torch

    from nebula.sim import coupled

    ic = workload(args.n)
    cfg = config(args.precision)
    learned = args.precision == "mixed" and not args.no_learned
    if learned:  # the headline's learned start: trained on 3 logged steps, then online
        _, _, _, _, info = run_gpu(ic, cfg, 1, 1, learned=True)
        cfg = info["_cfg"]
    st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg)
    sp = None
    if args.precision == "mixed":  # the headline's driver
        from nebula.sim.stepper import PipelinedStepper

        sp = PipelinedStepper(st, cfg)
    windows = []
    done = 0
    while done < args.long_run:
        w = min(50, args.long_run - done)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(w):
            if sp is not None:
                sp.step()
                st = sp.state
            else:
                st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
            its.append(st.iterations.double().mean())
        torch.cuda.synchronize()
        iters = float(torch.stack(its).sum().item())
        done += w
        free, total = torch.cuda.mem_get_info()
        rec = {"steps": done, "t": round(st.t, 3), "ms_per_step": round((time.perf_counter() - t0) / w * 1e3, 3),
               "newton_iters": round(iters / w, 4), "search_entries_per_particle": round(st.n_search_entries / args.n, 1),
               "torch_allocated_GB": round(torch.cuda.memory_allocated() / 1e9, 2),
               "device_used_GB": round((total - free) / 1e9, 2)}  # fmt: skip
        windows.append(rec)
        print(f"bench: {json.dumps(rec)}", file=sys.stderr, flush=True)
    print(json.dumps({"workload": f"stream crash, {args.n} particles", "precision": args.precision,
                      "learned_h": learned, "driver": "pipelined" if sp is not None else "eager",
                      "fallbacks": None if sp is None else sp.fallbacks, "windows": windows}))


def _per_step(d, steps):
    return {k: round(v[1] / steps, 4) for k, v in sorted(d.items(), key=lambda kv: -kv[1][1])}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (sharding the particles "
                                   "across GPUs is designed, not implemented)"}))  # fmt: skip
        sys.exit(2)
    from nebula import _hip

    dev = _hip.torch_device()
    if args.ensemble:
        run_ensemble(args)
        return
    if args.long_run > 0:
        run_long(args)
        return
    if args.report:
        run_report(args)
        return
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8, "launches": 5}))
        return
    ic = workload(args.n)
    cfg = config(args.precision)
    n = args.n
    print(f"bench: {n} particles, {args.precision}, warmup {args.warmup}, steps {args.steps}", file=sys.stderr,
          flush=True)  # fmt: skip
    learned = args.precision == "mixed" and not args.no_learned
    driver = "pipelined" if args.precision == "mixed" else "eager"
    wall, kern, entry, st, info = run_gpu(ic, cfg, args.steps, args.warmup, learned=learned, driver=driver)
    value = n * args.steps / wall
    print(f"bench: {value:.4g} particle-steps/s", file=sys.stderr, flush=True)
    # algorithmic work of the pair kernels, from this run's own lists and counters
    n_search = st.n_search_entries / n
    n_pairs = (st.n_tight_entries - n) / n
    k_it = float(st.iterations.double().mean().item())
    dens = (FLOPS_SEARCH_ENTRY * n_search + FLOPS_NEWTON_ENTRY * (k_it + 1) * n_search) * n
    fp64 = args.precision == "fp64"
    # kernel -> (bound, work per launch, unit)
    models = {
        # neighbour gather (+ the first Newton pass): write each search entry's index (the list
        # write of the neighbour build), read each particle's (x, y, z, h) record and mass once and
        # write its FP32 mass; the bit-exact layout also writes each entry's FP64 d2 and mass
        "k_gather_mixed": ("hbm", 4 * n_search * n + (32 + 8 + 4) * n, "GB/s"),
        "k_gather_solve": ("hbm", (4 + 8 + 8) * n_search * n + 32 * n, "GB/s"),
        "k_forces64": ("mfma", FLOPS_FORCE_PAIR * n_pairs * n, "TFLOP/s"),
        "k_forces_mixed": ("mfma", FLOPS_FORCE_PAIR * n_pairs * n, "TFLOP/s"),
        "k_gradients64": ("mfma", FLOPS_GRAD_PAIR * n_pairs * n, "TFLOP/s"),
        "k_gradients_mixed": ("mfma", FLOPS_GRAD_PAIR * n_pairs * n, "TFLOP/s"),
        # tightening: read each search entry (index; the bit-exact layout also FP64 d2 and the
        # neighbour's h), each particle's (x, y, z, h) record and FP32 mass once, write the kept index
        "k_tight_emit": ("hbm", (4 + 8 + 8) * n_search * n + 4 * (n_pairs + 1) * n, "GB/s"),
        "k_tight_emit_rm": ("hbm", 4 * n_search * n + (32 + 4) * n + 4 * (n_pairs + 1) * n, "GB/s"),
        # the FP64 step on the lane-group lists: the Newton solve over each stored row (its index
        # read once per evaluation, the records from cache), and the tightening of the row merged
        # with its extras into the 64-row blocks
        "k_solve64": ("mfma", FLOPS_NEWTON_ENTRY * (k_it + 1) * (n_pairs + 1) * n, "TFLOP/s"),
        "k_emit64": ("hbm", 4 * n_search * n + 32 * n + 4 * (n_pairs + 1) * n, "GB/s"),
    }
    timed = {k: v for k, v in kern.items() if k in models}
    dom = max(timed, key=lambda k: timed[k][1])
    launches, total_ms = timed[dom]
    avg_s = total_ms / launches / 1e3
    fpeak = PEAK["fp64"] if fp64 else PEAK["fp32"]
    flops = step_flops(st, n, n_search, n_pairs, k_it, value, kern, args.steps, fp64)
    bound, work, unit = models[dom]
    if dom in flops["kernel_flop_per_particle"]:
        # the gather and emission are priced as the compute they do (their PMC counts show vector
        # issue, not HBM, bounding them); their HBM figure stays alongside
        hbm_work = work
        bound, work, unit = "mfma", flops["kernel_flop_per_particle"][dom] * n, "TFLOP/s"
    else:
        hbm_work = None
    if bound == "hbm":
        achieved, peak = work / avg_s / 1e9, HBM_PEAK
    else:
        achieved, peak = work / avg_s / 1e12, fpeak
    out = {
        "metric": "SPH particle-steps/s at 1M particles (plain FP64 vs mixed+learned-h; % FP32 peak)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else ("f32+f64+bf16" if learned else "f32+f64"),
        "data": "synthetic (seeded stream-crash initial conditions)",
        "config": {"workload": "stream crash, 2 x 524,288-particle streams (1,048,576)" if n == 1_048_576 else
                   f"stream crash, {n} particles", "precision": args.precision, "kernel": "cubic, eta 1.2",
                   "viscosity": "Monaghan alpha 1 beta 2 + Balsara", "cooling": "power law rho^2 T^0.5",
                   "driver": driver},
        "roofline": {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": _traffic(dom, args.precision),
                     "avg_launch_ms": avg_s * 1e3,
                     "note": "compute-bound pair arithmetic runs on the vector ALUs and is priced against "
                             "MI355X's dense FP64 (78.6) or FP32 (157.3) TFLOP/s, which the vector and matrix "
                             "units share",
                     "work": f"{work / n:.0f} {'FLOP' if bound != 'hbm' else 'B'}/particle per launch "
                             f"(n_s {n_search:.1f}, pairs {n_pairs:.1f}, Newton iterations {k_it:.2f})",
                     **_issue_and_hbm(dom, args.precision, avg_s, hbm_work)},
        "fp32_peak": flops["report"],
        "kernels_ms_per_step": _per_step(kern, args.steps),
        "gather_launches_per_step": kern.get("k_gather_mixed", kern.get("k_gather_solve", (0, 0.0)))[0] / args.steps,
        "learned_h": {k: v for k, v in info.items() if not k.startswith("_")} if learned else None,
        "newton_iters": info["newton_iters"],
    }
    if learned:
        print("bench: per-field parity (mixed vs FP64, learned vs plain start)", file=sys.stderr, flush=True)
        out["parity"] = parity_report(st, info["_cfg"])
    if not args.no_compare:
        other = "fp64" if args.precision == "mixed" else "mixed"
        del st
        out["compare"] = {}
        # the other precision (eager steps), the mixed mode without the learned start (same driver),
        # and the headline's own configuration stepped eagerly (coupled.stable_dt + coupled.step)
        alongside = [(other, other, False, "eager")]
        if learned:
            alongside += [("mixed", "mixed", False, driver), ("mixed+learned, stepper", "mixed", True, "stepper"),
                          ("mixed+learned, eager steps", "mixed", True, "eager")]  # fmt: skip
        for label, name, lrn, drv in alongside:
            print(f"bench: {label} alongside", file=sys.stderr, flush=True)
            w2, k2, _, _, info2 = run_gpu(ic, config(name), args.steps, args.warmup, learned=lrn, driver=drv)
            v2 = n * args.steps / w2
            gk = k2.get("k_gather_mixed", k2.get("k_gather_solve", (0, 0.0)))
            out["compare"][label] = {"value": v2, "ms_per_step": w2 / args.steps * 1e3, "value_ratio": value / v2,
                                     "driver": drv, "newton_iters": info2["newton_iters"],
                                     "gather_launches_per_step": gk[0] / args.steps,
                                     "kernels_ms_per_step": _per_step(k2, args.steps)}  # fmt: skip
    if args.drift_steps > 0:
        runs = {args.precision + ("+learned" if learned else ""): info["_cfg"] if learned else cfg}
        if not args.no_compare:
            runs[other] = config(other)
        out["drift"] = {}
        for name, c in runs.items():
            print(f"bench: {args.drift_steps}-step ledger run, {name}", file=sys.stderr, flush=True)
            out["drift"][name] = drift_run(ic, c, args.drift_steps)
        out["drift"]["tolerance"] = "max |dE| / sum m|H| < 1e-5 (the reference's coupled ledger tests)"
    if args.run_steps > 0:
        out["whole_run"] = whole_run(ic, info["_cfg"] if learned else cfg, args.run_steps, driver,
                                     None if (args.no_compare or fp64) else config("fp64"))
    if not args.no_cpu:
        print("bench: CPU baseline (FP64 oracle, OpenMP)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(ic, cfg, args.cpu_steps)
    print(json.dumps(out))


def step_flops(st, n, n_search, n_pairs, k_it, value, kern, steps, fp64):
    """Algorithmic FLOPs of one particle-step, counted from the reference's source operations
    (FMA = 2, div = sqrt = 1): 10 n_s (distance pass) + 20 (k + 1) n_d (Newton sums and the final
    density) + 83 n_f (force pair) + 45 n_f (gradient pair: the Balsara limiter) + 35 (i_1 + i_2 + 2)
    (GLF) + 40 (kicks, EOS, cooling). n_s, n_f and k come from this run's lists and counters;
    n_d, the entries inside 2 h_search of the row, is taken as n_f + 1 (the tight rows); i_1 + i_2
    is the GLF's own iteration count on the final state. Also the FLOPs of each pair kernel per
    launch (its share of that sum) and its TFLOP/s from the kernel timing."""
    from nebula.sim import coupled
    from nebula.sim.integrator import glf_device

    dt = coupled.stable_dt(st, config("fp64" if fp64 else "mixed"))
    _, _, it = glf_device(st.x, st.p, dt, raise_on_failure=False)
    i12 = float(it.clamp(min=0).double().mean().item())
    n_d = n_pairs + 1.0
    terms = {"distance": FLOPS_SEARCH_ENTRY * n_search, "newton": FLOPS_NEWTON_ENTRY * (k_it + 1.0) * n_d,
             "forces": FLOPS_FORCE_PAIR * n_pairs, "gradients": FLOPS_GRAD_PAIR * n_pairs,
             "glf": FLOPS_GLF_ITER * (i12 + 2.0), "per_particle": float(FLOPS_PER_PARTICLE)}  # fmt: skip
    per = sum(terms.values())
    peak = PEAK["fp64"] if fp64 else PEAK["fp32"]
    # each kernel's share: the gather runs the distance pass and the first Newton evaluation, the
    # emission the final density sums (and the tightening's distances again, not counted twice)
    kernel_flop = {
        # mixed: the first Newton evaluations run in the gather; FP64: all of them in k_solve64
        "k_gather_mixed": FLOPS_SEARCH_ENTRY * n_search + (0.0 if fp64 else FLOPS_NEWTON_ENTRY * k_it * n_d),
        "k_solve64": FLOPS_NEWTON_ENTRY * (k_it + 1.0) * n_d,
        "k_tight_emit_rm": FLOPS_NEWTON_ENTRY * n_d,
        "k_forces_mixed": FLOPS_FORCE_PAIR * n_pairs,
        "k_divcurl_mixed": FLOPS_GRAD_PAIR * n_pairs,
        "k_gather_solve": FLOPS_SEARCH_ENTRY * n_search + FLOPS_NEWTON_ENTRY * (k_it + 1.0) * n_d,
        "k_forces64": FLOPS_FORCE_PAIR * n_pairs,
        "k_gradients64": FLOPS_GRAD_PAIR * n_pairs,
    }
    kernels = {}
    for name, f in kernel_flop.items():
        if name in kern and kern[name][0]:
            avg_s = kern[name][1] / kern[name][0] / 1e3
            tf = f * n / avg_s / 1e12
            kernels[name] = {"flop_per_particle": round(f, 1), "avg_launch_ms": round(avg_s * 1e3, 4),
                             "tflops": round(tf, 3), "frac_of_peak": round(tf / peak, 4)}  # fmt: skip
    step_tf = per * value / 1e12
    report = {"peak_tflops": peak, "precision_peak": "fp64" if fp64 else "fp32",
              "step": {"flop_per_particle_step": round(per, 1), "terms": {k: round(v, 1) for k, v in terms.items()},
                       "tflops": round(step_tf, 3), "frac_of_peak": round(step_tf / peak, 4),
                       "glf_iterations": round(i12, 3)},
              "kernels": kernels}  # fmt: skip
    return {"report": report, "kernel_flop_per_particle": {k: v for k, v in kernel_flop.items() if k in kernels}}


def _issue_and_hbm(kernel, precision, avg_s, hbm_work):
    """Second views of the dominant kernel: its vector-instruction issue (SQ_INSTS_VALU from the
    committed PMC summary, at one wave64 instruction per 4 cycles per SIMD) as a fraction of its
    launch, and its algorithmic HBM bytes over the launch against 8 TB/s."""
    out = {}
    f = REPO / "profiles" / "traffic.json"
    rec = json.loads(f.read_text()).get(precision, {}).get(kernel, {}) if f.exists() else {}
    insts = rec.get("valu_insts_per_launch")
    if insts:
        issue_s = insts / VALU_ISSUE_PEAK
        out["valu_issue"] = {"insts_per_launch": insts, "issue_ms": round(issue_s * 1e3, 4),
                             "frac_of_launch": round(issue_s / avg_s, 4)}  # fmt: skip
    if hbm_work is not None:
        gbs = hbm_work / avg_s / 1e9
        out["hbm_view"] = {"algorithmic_bytes_per_launch": int(hbm_work), "achieved_gbs": round(gbs, 1),
                           "frac_of_8tbs": round(gbs / HBM_PEAK, 4)}  # fmt: skip
    return out


def whole_run(ic, cfg, steps, driver, cfg64):
    """The workload's whole run: ``steps`` steps from the initial state (the shocked layer grows
    and the step gets dearer), timed as a whole: mean ms per step and particle-steps/s, for the
    headline mode and, when given, the FP64 path (eager steps) over the same number of steps."""
    import torch

    from nebula.sim import coupled

    out = {"steps": steps}
    n = len(ic.x)
    runs = [("headline", cfg, driver)] + ([("fp64", cfg64, "eager")] if cfg64 is not None else [])
    for label, c, drv in runs:
        print(f"bench: whole {steps}-step run, {label}", file=sys.stderr, flush=True)
        st = coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if drv in ("stepper", "pipelined"):
            from nebula.sim.stepper import PipelinedStepper, Stepper

            sp = (PipelinedStepper if drv == "pipelined" else Stepper)(st, c)
            for _ in range(steps):
                sp.step()
            extra = {"stepper_fallbacks": sp.fallbacks}
        else:
            for _ in range(steps):
                st = coupled.step(st, c, coupled.stable_dt(st, c))
            extra = {}
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[label] = {"ms_per_step": wall / steps * 1e3, "value": n * steps / wall, "driver": drv, **extra}
    if "fp64" in out:
        out["value_ratio"] = out["headline"]["value"] / out["fp64"]["value"]
    return out


def _traffic(kernel, precision):
    """HBM bytes per launch of ``kernel`` from the committed PMC summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), or None when it has not been measured."""
    f = REPO / "profiles" / "traffic.json"
    if not f.exists():
        return None
    data = json.loads(f.read_text())
    rec = data.get(precision, {}).get(kernel)
    return None if rec is None else rec.get("bytes_per_launch")


if __name__ == "__main__":
    main()
