"""One cloud stepped with a single host synchronisation per step (no counterpart in the
reference, whose driver calls ``stable_dt`` and ``step`` in turn).

``coupled.step`` sizes its list buffers from the data as it goes: the density pass reads its
row totals back before it allocates, the GN flow and the equation of state are checked on the
host, and ``stable_dt`` reads the next step's minima. Each of those reads drains the GPU's queue,
and the GPU then waits while the host prepares the next launches.

``Stepper`` runs the same step with the fixed-capacity passes of a recorded ensemble step
(``nb_fixed_reserve``; capacities from the previous step's row counts, with headroom). What does
not fit, and what the eager step would raise on, is counted in a device status word instead. The
step ends with the next step's time-step minima on the device. The host reads the status, the
minima and two list totals in one copy, the step's only synchronisation. A step with a nonzero
status is discarded and done again by ``coupled.step`` (which grows its buffers, rebuilds its
search, or raises the reference's exceptions); the capacities are then taken from that step.

The kernels and their arguments are those of ``coupled.step`` in the mixed-precision mode, so the
states are bit-identical (``tests/test_gpu_stepper.py``), the learned start and its online
training included: its pairs are logged after the status check, in the eager step's order.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import replace

import numpy as np
import torch

from .. import _hip
from . import coupled
from .captured import STATUS_NAMES, _Fixed, ctypes_void


class Stepper:
    """Steps a single-cloud mixed-precision ``GasState`` (from coupled.initialise or a step).

    >>> sp = Stepper(st, cfg)
    >>> for _ in range(1000):
    ...     sp.step()                    # returns the dt taken (coupled.stable_dt of the state)
    >>> st = sp.state
    """

    def __init__(self, st: coupled.GasState, cfg: coupled.CoupledConfig, headroom: float = 1.3):
        hc = cfg.hydro
        if hc.precision != "mixed":
            raise ValueError("the stepper runs the mixed-precision passes")
        if st.cloud is not None:
            raise ValueError("ensembles: nebula.sim.captured.CapturedEnsemble")
        if st.row_counts is None:
            raise ValueError("the state needs the row counts of a density pass (coupled.initialise or step)")
        self.cfg = cfg
        self.state = st
        self.headroom = float(headroom)
        n = int(st.x.shape[0])
        dev = st.x.device
        self.n = n
        self.cloud = torch.zeros(n, dtype=torch.int32, device=dev)
        self.seg_off_dev = torch.tensor([0, n], dtype=torch.int64, device=dev)
        self.status = torch.zeros(16, dtype=torch.int32, device=dev)
        self.minima = torch.empty(4, dtype=torch.float64, device=dev)
        self.steps = 0
        self.fallbacks = 0
        self.fallback_reasons = []
        self._reserve()
        self.dt = coupled.stable_dt(st, cfg)

    def _reserve(self):
        """Capacities of the fixed passes from the state's row counts, as the library sizes rows
        from a hint (a quarter + 32 entries per gather row, extras doubled + 16, multiples of 4),
        times the headroom."""
        st = self.state
        L = _hip.lib()
        sizes = np.zeros(6, np.int64)
        _hip.check(L.nb_ctx_sizes(_hip.ctx(), sizes.ctypes.data_as(_hip._P64)), "nb_ctx_sizes")
        c = st.row_counts.cpu().numpy().astype(np.int64)
        a, b = c[:, 0], c[:, 1]
        need = (int(np.sum((a + a // 4 + 32 + 3) // 4 * 4)), int(np.sum((2 * b + 16 + 3) // 4 * 4)), int(np.sum(a + b)),
                max(int(sizes[3]), 1))  # fmt: skip
        gather, extras, index, cells = (int(v * self.headroom) + 4096 for v in need)
        self.capacities = {"gather": gather, "extras": extras, "index": index, "cells": cells}
        _hip.check(L.nb_fixed_reserve(_hip.ctx(), self.n, 1, gather, extras, index, cells,
                                      int(not (st.m_uniform > 0.0))), "nb_fixed_reserve")  # fmt: skip
        self.index = torch.empty(index, dtype=torch.int32, device=st.x.device)
        self.fixed = _Fixed(self.status, self.seg_off_dev, 1, self.index, cloud=self.cloud)
        self._generation = _hip.generation()

    def step(self) -> float:
        """One step of coupled.stable_dt(state); returns the dt taken."""
        st, cfg, dt = self.state, self.cfg, self.dt
        if _hip.generation() != self._generation:
            self._reserve()  # the library's scratch was released or moved since the reservation
        gate = _gate(cfg)
        gate_before = None if gate is None else gate.snapshot()
        self.status.zero_()
        self.fixed.pairs = None
        new = coupled.step(st, cfg, dt, fixed=self.fixed)
        _hip.call("nb_dt_minima_dev", _hip.ptr(new.h), _hip.ptr(new.vsig_max), _hip.ptr(new.a_hydro),
                  _hip.ptr(new.nu), 0.25, _hip.ptr(new.x), self.n, _hip.ptr(self.minima))  # fmt: skip
        nb_len = new.nneigh
        totals = torch.stack([new.row_counts.sum(dtype=torch.int64), nb_len.sum().to(torch.int64)]).double()
        packed = torch.cat([self.status.double(), self.minima, totals]).cpu().numpy()  # the one synchronisation
        status = packed[:16].astype(np.int64)
        if np.any(status[:9] != 0):
            self.fallbacks += 1
            self.fallback_reasons.append({STATUS_NAMES[k]: int(status[k]) for k in range(9) if status[k]})
            if gate is not None:  # the step is done again: the same decision, observed once
                gate.restore(gate_before)
            new = coupled.step(st, cfg, dt)
            self.state = new
            self.headroom = max(self.headroom, 1.3)
            self._reserve()
            self.dt = coupled.stable_dt(new, cfg)
        else:
            model, feat, h_prev, h_final = self.fixed.pairs
            coupled.log_pairs(cfg, model, feat, h_prev, h_final)
            self.state = replace(new, n_search_entries=int(packed[20]), n_tight_entries=int(packed[21]))
            self.dt = coupled.dt_from_minima(packed[16:20], cfg, new.nu is not None)
        self.steps += 1
        return dt


class PipelinedStepper(Stepper):
    """``Stepper`` that launches step k + 1 before it reads step k back.

    ``Stepper`` ends each step with its one host synchronisation, and the GPU then waits while the
    host launches the next step's first, short kernels one by one. Here the next step is launched
    first, so the host prepares step k + 1 while the GPU still runs step k:

    - each step's dt is taken on the device from the previous step's minima (nb_dt_advance) and
      read from device memory by the step's kernels, as in a recorded step;
    - the host then reads step k's status, minima and totals (copied to pinned memory when step k
      ended), computes dt from the minima as ``coupled.dt_from_minima`` does, and keeps step k + 1
      only when step k was good and the device took the same dt bit for bit; otherwise step k + 1
      is discarded (and, for a bad step k, step k done again by ``coupled.step``, as ``Stepper``
      does);
    - the learned start's gate decides from the counts of every step but the last, so its
      decision for step k + 1 needs nothing of step k; a step that runs the network waits for
      step k's pairs (its online training comes first), so it is launched after the read-back.

    States are those of ``coupled.step`` with ``coupled.stable_dt``, bit for bit
    (``tests/test_gpu_stepper.py``). ``state`` is the last step read back; one more may be running.
    Covers the physics whose dt a step can take from the device: power-law cooling, constant or
    Balsara viscosity, GN or Newtonian flow, with or without the learned start.
    """

    def __init__(self, st: coupled.GasState, cfg: coupled.CoupledConfig, headroom: float = 1.3):
        hc = cfg.hydro
        if math.isfinite(cfg.cooling_beta) or (hc.viscosity and hc.switch):
            raise ValueError("the pipelined step takes its dt from the device: power-law cooling and the constant "
                             "or Balsara viscosity only")  # fmt: skip
        if hc.ns_nu > 0 or hc.ns_alpha > 0 or hc.disc_alpha_av > 0:
            raise ValueError("the pipelined step covers the cloud physics without disc viscosity")
        self.dts = torch.zeros(2, dtype=torch.float64, device=st.x.device)
        super().__init__(st, cfg, headroom)
        self._pinned = [torch.empty(24, dtype=torch.float64).pin_memory() for _ in range(2)]
        # nb_step_readback's output (the 24 values read back, then the gate's slow-row count) and
        # its three counters (zero between calls)
        self._rb_out = torch.zeros(25, dtype=torch.float64, device=st.x.device)
        self._rb_acc = torch.zeros(3, dtype=torch.int64, device=st.x.device)
        self._slot = 0
        self._flight = None  # the launched step not yet read back
        self.overlapped = 0  # steps launched before their predecessor was read back
        self.discarded = 0  # launched steps thrown away (their predecessor failed, or its dt differed)

    def _reserve(self):
        super()._reserve()
        self.fixed.dts = self.dts
        self.fixed.reads_back = True  # the gate's count comes from nb_step_readback

    def _launch(self, st, host_dt=None):
        """Enqueue one step from st. host_dt: its dt from the host (written to the device first);
        None: the dt the previous launch left on the device."""
        cfg = self.cfg
        gate = _gate(cfg)
        snap = None if gate is None else gate.snapshot()
        self.fixed.use_net = bool(cfg.hydro.learned_h and (gate is None or gate.decide()))
        if host_dt is not None:
            self.dts[0].fill_(float(host_dt))
            self.dts[1].fill_(float(st.dt_prev))
            self.status.zero_()  # later launches find it cleared by nb_step_readback
        self.fixed.pairs = None
        # coupled.step takes the dt from self.dts; the host value only dates the state (set again
        # once the step is read back)
        new = coupled.step(st, cfg, self.dt if host_dt is None else host_dt, fixed=self.fixed)
        _hip.call("nb_dt_minima_dev", _hip.ptr(new.h), _hip.ptr(new.vsig_max), _hip.ptr(new.a_hydro),
                  _hip.ptr(new.nu), 0.25, _hip.ptr(new.x), self.n, _hip.ptr(self.minima))  # fmt: skip
        hc = cfg.hydro
        _hip.call("nb_dt_advance", _hip.ptr(self.minima), float(hc.courant), float(hc.c_force),
                  int(new.nu is not None), int(bool(cfg.hole)), float(cfg.eta_orb), _hip.ptr(self.dts))  # fmt: skip
        if new.iterations.dtype != torch.int32 or new.row_counts.dtype != torch.int32:
            raise TypeError("nb_step_readback takes int32 iteration and row counts")
        _hip.call("nb_step_readback", _hip.ptr(self.status), int(self.status.numel()), _hip.ptr(self.minima),
                  _hip.ptr(new.row_counts), _hip.ptr(new.nneigh), _hip.ptr(new.iterations), self.n,
                  _hip.ptr(self.dts), _hip.ptr(self._rb_acc), _hip.ptr(self._rb_out))  # fmt: skip
        if gate is not None:
            gate.observe(self._rb_out[24], self.n, self.fixed.use_net)
        buf = self._pinned[self._slot]
        self._slot ^= 1
        buf.copy_(self._rb_out[:24], non_blocking=True)
        done = torch.cuda.Event()
        done.record()
        return {"state": new, "buf": buf, "done": done, "snap": snap, "pairs": self.fixed.pairs}

    def _next_uses_net(self) -> bool:
        cfg = self.cfg
        if not cfg.hydro.learned_h:
            return False
        gate = _gate(cfg)
        if gate is None:
            return True
        snap = gate.snapshot()
        use = gate.decide()
        gate.restore(snap)
        return use

    def step(self) -> float:
        """One step of coupled.stable_dt(state) (read back); returns the dt taken."""
        cfg, dt = self.cfg, self.dt
        if self._flight is None:
            if _hip.generation() != self._generation:
                self._reserve()  # the library's scratch was released or moved since the reservation
            self._flight = self._launch(self.state, host_dt=dt)
        cur, nxt = self._flight, None
        if not self._next_uses_net():
            nxt = self._launch(cur["state"])
            self.overlapped += 1
        cur["done"].synchronize()
        packed = cur["buf"].numpy().copy()
        status = packed[:16].astype(np.int64)
        gate = _gate(cfg)
        if np.any(status[:9] != 0):
            self.fallbacks += 1
            self.fallback_reasons.append({STATUS_NAMES[k]: int(status[k]) for k in range(9) if status[k]})
            if gate is not None:  # undoes both launches' decisions; the step is decided again
                gate.restore(cur["snap"])
            if nxt is not None:
                self.discarded += 1
            torch.cuda.synchronize()  # the discarded launch has finished with the shared buffers
            new = coupled.step(self.state, cfg, dt)
            self.state = new
            self.headroom = max(self.headroom, 1.3)
            self._reserve()
            self.dt = coupled.stable_dt(new, cfg)
            self._flight = None
            self.steps += 1
            return dt
        model, feat, h_prev, h_final = cur["pairs"]
        coupled.log_pairs(cfg, model, feat, h_prev, h_final)
        new = cur["state"]
        self.state = replace(new, t=self.state.t + dt, dt_prev=float(dt), n_search_entries=int(packed[20]),
                             n_tight_entries=int(packed[21]))  # fmt: skip
        self.dt = coupled.dt_from_minima(packed[16:20], cfg, new.nu is not None)
        if nxt is not None and packed[22] != self.dt:  # the device's dt differs from the host's
            if gate is not None:
                gate.restore(nxt["snap"])
            nxt = None
            self.discarded += 1
            torch.cuda.synchronize()
        self._flight = nxt
        self.steps += 1
        return dt


def _gate(cfg):
    """The learned start's gate of the configuration, or None."""
    model = cfg.hydro.learned_model if cfg.hydro.learned_h else None
    return getattr(model, "gate", None)


# the state fields a recorded step carries from one replay to the next (those the state has)
_FIELDS = ("x", "p", "u", "mass", "h", "alpha", "a_hydro", "dudt", "rho", "vsig_max", "c_s", "omega", "iterations",
           "flags", "perm", "row_counts", "div_v", "curl_v", "nneigh", "e_radiated_dev")  # fmt: skip


class RecordedStepper(Stepper):
    """``Stepper`` whose step is recorded once as a hipGraph and replayed.

    The recorded step is ``coupled.step`` through the fixed-capacity passes, with its dt (and the
    previous dt the learned start's features use) read from device memory: the host writes them
    before each replay. Each replay saves the state it starts from, runs the step, copies the new
    state over the old and leaves the status word and the next dt's minima on the device; the
    host reads them in one copy, as ``Stepper`` does. A replay with a nonzero status is undone
    (the saved state restored), the step done by ``coupled.step`` and the graph recorded again.
    The learned start's online training runs between replays, on the pairs of the replay.
    The graph holds the addresses of the library's scratch buffers; when an eager call since the
    recording has grown (moved) one of them (``nb_ctx_generation``), the step is recorded again
    before the next replay.

    Covers the physics a recorded step can take its dt from the device for: GN or Newtonian flow,
    power-law cooling, constant or Balsara viscosity, the learned start (no beta cooling, no
    Cullen-Dehnen switch, no disc viscosity). Replays are bit-identical to ``coupled.step``
    (``tests/test_gpu_stepper.py``); the graph holds no memset node.
    """

    def __init__(self, st: coupled.GasState, cfg: coupled.CoupledConfig, headroom: float = 1.3):
        hc = cfg.hydro
        if hc.ns_nu > 0 or hc.ns_alpha > 0 or hc.disc_alpha_av > 0:
            raise ValueError("the recorded step covers the cloud physics without disc viscosity")
        super().__init__(st, cfg, headroom)
        self.replays = 0
        self.rerecords = 0  # recordings redone because library scratch moved under them
        self._record()

    def _record(self):
        st = self.state
        dev = st.x.device
        self.fields = tuple(k for k in _FIELDS if getattr(st, k) is not None)
        self.S = {k: getattr(st, k).clone() for k in self.fields}
        self.backup = {k: torch.empty_like(v) for k, v in self.S.items()}
        self.dts = torch.tensor([self.dt, st.dt_prev], dtype=torch.float64, device=dev)
        self.fixed.dts = self.dts
        self._template = st
        self._host_dts = torch.empty(2, dtype=torch.float64).pin_memory()
        # one pass of the recorded sequence outside the capture (first-call set-up), then the capture
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._body(commit=False)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph(keep_graph=True)
        with torch.cuda.graph(self.graph):
            self._body(commit=True)
        self.graph.instantiate()
        torch.cuda.synchronize()
        counts = np.zeros(16, np.int32)
        _hip.check(_hip.lib().nb_graph_node_counts(ctypes_void(self.graph.raw_cuda_graph()),
                                                   counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                   "nb_graph_node_counts")  # fmt: skip
        self.node_counts = {"kernel": int(counts[0]), "memcpy": int(counts[1]), "memset": int(counts[2])}
        if counts[2]:
            raise RuntimeError("the recorded step holds memset nodes; it is not replayed")
        self._generation = _hip.generation()

    def _body(self, commit: bool):
        """The recorded step: S -> S (S saved in backup first), status and minima written."""
        S = self.S
        for k, v in S.items():
            self.backup[k].copy_(v)
        self.status.zero_()
        self.fixed.pairs = None
        st = replace(self._template, **{k: S[k] for k in self.fields})
        new = coupled.step(st, self.cfg, self.dt, fixed=self.fixed)
        _hip.call("nb_dt_minima_dev", _hip.ptr(new.h), _hip.ptr(new.vsig_max), _hip.ptr(new.a_hydro),
                  _hip.ptr(new.nu), 0.25, _hip.ptr(new.x), self.n, _hip.ptr(self.minima))  # fmt: skip
        self._pairs = self.fixed.pairs
        self._fallbacks_dev = new.learned_fallbacks_dev
        if commit:
            for k in self.fields:
                S[k].copy_(getattr(new, k))
        else:
            for k in self.fields:  # the untimed pass leaves the state as it was
                S[k].copy_(self.backup[k])

    def step(self) -> float:
        """One replayed step of coupled.stable_dt(state); returns the dt taken."""
        st, cfg, dt = self.state, self.cfg, self.dt
        if _hip.generation() != self._generation:
            # an eager call since the recording grew (moved) or released library scratch the graph
            # points at: reserve the fixed passes' buffers again and record anew
            self.rerecords += 1
            self._reserve()
            self._record()
        self._host_dts[0], self._host_dts[1] = dt, st.dt_prev
        self.dts.copy_(self._host_dts, non_blocking=True)
        self.graph.replay()
        S = self.S
        totals = torch.stack([S["row_counts"].sum(dtype=torch.int64), S["nneigh"].sum().to(torch.int64)]).double()
        packed = torch.cat([self.status.double(), self.minima, totals]).cpu().numpy()  # the one synchronisation
        status = packed[:16].astype(np.int64)
        self.replays += 1
        if np.any(status[:9] != 0):
            self.fallbacks += 1
            self.fallback_reasons.append({STATUS_NAMES[k]: int(status[k]) for k in range(9) if status[k]})
            for k, v in self.backup.items():
                S[k].copy_(v)
            new = coupled.step(st, cfg, dt)
            self.state = new
            self.headroom = max(self.headroom, 1.3)
            self._reserve()
            self.dt = coupled.stable_dt(new, cfg)
            self._record()
        else:
            model, feat, h_prev, h_final = self._pairs
            coupled.log_pairs(cfg, model, feat, h_prev, h_final)
            self.state = replace(self._template, t=st.t + dt, dt_prev=float(dt),
                                 n_search_entries=int(packed[20]), n_tight_entries=int(packed[21]),
                                 learned_fallbacks_dev=self._fallbacks_dev,
                                 **{k: S[k] for k in self.fields})  # fmt: skip
            self.dt = coupled.dt_from_minima(packed[16:20], cfg, st.nu is not None)
        self.steps += 1
        return dt

    def sync_back(self) -> coupled.GasState:
        """The replayed state as a GasState of its own (copies of the recorded step's buffers)."""
        self.state = replace(self.state, **{k: v.clone() for k, v in self.S.items()})
        return self.state
