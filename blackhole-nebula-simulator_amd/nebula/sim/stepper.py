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

from dataclasses import replace

import numpy as np
import torch

from .. import _hip
from . import coupled
from .captured import STATUS_NAMES, _Fixed


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

    def step(self) -> float:
        """One step of coupled.stable_dt(state); returns the dt taken."""
        st, cfg, dt = self.state, self.cfg, self.dt
        self.status.zero_()
        self.fixed.pairs = None
        new = coupled.step(st, cfg, dt, fixed=self.fixed)
        _hip.call("nb_dt_minima_dev", _hip.ptr(new.h), _hip.ptr(new.vsig_max), _hip.ptr(new.a_hydro),
                  _hip.ptr(new.nu), 0.25, _hip.ptr(new.x), self.n, _hip.ptr(self.minima))  # fmt: skip
        nb_len = new.nneigh
        totals = torch.stack([new.row_counts.sum(dtype=torch.int64), nb_len.sum().to(torch.int64)]).double()
        packed = torch.cat([self.status.double(), self.minima, totals]).cpu().numpy()  # the one synchronisation
        status = packed[:16].astype(np.int64)
        if np.any(status[:8] != 0):
            self.fallbacks += 1
            self.fallback_reasons.append({STATUS_NAMES[k]: int(status[k]) for k in range(8) if status[k]})
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
