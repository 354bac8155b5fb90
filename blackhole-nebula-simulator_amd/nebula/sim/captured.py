"""Ensemble steps recorded once as a hipGraph and replayed (no counterpart in the reference).

``CapturedEnsemble`` records one ``ensemble.step`` of every cloud (kicks, GN flow, grid order,
permutation, density and tight lists, equation of state, limiter, forces, kicks, and the next
step's time-step minima) into one graph, and replays it: one graph launch per step, no host
synchronisation inside it. The passes that size their buffers from the data run with capacities
fixed beforehand (``nb_fixed_reserve``, from an eager step with headroom) and report in a device
status word what did not fit. The host reads the status words and the per-cloud time-step minima
after each replay (one small copy), computes each cloud's dt as ``ensemble.stable_dt_clouds``
does, and writes it to the device for the next replay. If a replay reports anything, the state
saved at its start is restored, the step is done eagerly, and the graph is recorded again with
capacities from that step. The graph also holds the addresses of the library's scratch buffers:
when an eager call has grown (moved) one of them since the recording (``nb_ctx_generation``),
the step is recorded again before the next replay.

The graph's kernels and their arguments are those of the eager ensemble step, and the grids are
planned on the device with the host plan's arithmetic, so a replayed step gives the eager step's
bits (``tests/test_gpu_captured.py``). The recorded graph holds no memset node: every clear is a
kernel.
"""

from __future__ import annotations

import ctypes
from dataclasses import replace

import numpy as np
import torch

from .. import _dev, _hip
from . import coupled, ensemble
from .neighbours import RowLists, permute_many
from .sph import C_NORM, KERNEL_IDS, DensityResult

# the state fields a step carries from one replay to the next
_FIELDS = ("x", "p", "u", "mass", "h", "alpha", "a_hydro", "dudt", "rho", "vsig_max", "c_s", "omega", "iterations",
           "flags", "perm", "row_counts")  # fmt: skip
STATUS_NAMES = ("gather rows", "extras", "tight index", "grid cells", "h outgrew the search", "not in cell order",
                "GN flow failures", "u <= 0 or outside the EOS bracket", "EOS did not converge")  # fmt: skip


class _Fixed:
    """The fixed-capacity passes of a step with no host synchronisation (the ``fixed`` hook of
    coupled.step / coupled._hydro): grid order and density of the systems seg_off_dev delimits
    (cloud: the particles' system, or the state's own ``cloud``), what does not fit counted in
    ``status``."""

    def __init__(self, status, seg_off_dev, nseg, index, cloud=None):
        self.status = status
        self.seg_off_dev = seg_off_dev
        self.nseg = nseg
        self.index = index
        self.cloud = cloud
        self.pairs = None

    def order(self, x, h_search, cell_factor):
        n = int(x.shape[0])
        order = _dev.empty(n, torch.int32)
        _hip.call("nb_grid_order_seg_fixed", _hip.ptr(x), _hip.ptr(h_search), n, _hip.ptr(self.cloud),
                  _hip.ptr(self.seg_off_dev), self.nseg, float(cell_factor), _hip.ptr(self.status),
                  _hip.ptr(order))  # fmt: skip
        return order

    def density(self, st, cfg, h_start, counts):
        hc = cfg.hydro
        n = int(st.x.shape[0])
        kid = KERNEL_IDS[hc.kernel]
        h_guess = st.h.clone()
        h_search = coupled._search_lengths(cfg, h_guess)
        h, rho, omega = (_dev.empty(n) for _ in range(3))
        iters = _dev.empty(n, torch.int32)
        flags = _dev.empty(n, torch.int32)
        start = h_start if h_start is not None else h_guess
        _hip.call("nb_hydro_density_mixed_fixed", _hip.ptr(st.x), _hip.ptr(st.mass), float(st.m_uniform),
                  _hip.ptr(start), _hip.ptr(h_search), n, _hip.ptr(st.cloud if self.cloud is None else self.cloud),
                  _hip.ptr(self.seg_off_dev), self.nseg,
                  float(st.cell_factor), _hip.ptr(st.row_counts), _hip.ptr(counts), kid, float(C_NORM[kid]),
                  float(hc.eta), 1e-4, float(hc.h_max), _hip.ptr(h), _hip.ptr(rho), _hip.ptr(omega), _hip.ptr(iters),
                  _hip.ptr(flags), _hip.ptr(self.status))  # fmt: skip
        ln = _dev.empty(n, torch.int32)
        off = _dev.empty(n + 1, torch.int64)
        _hip.call("nb_hydro_tight_mixed_fixed", _hip.ptr(st.x), _hip.ptr(h), _hip.ptr(rho), _hip.ptr(omega),
                  _hip.ptr(off), _hip.ptr(ln), _hip.ptr(self.index), _hip.ptr(self.status))  # fmt: skip
        return DensityResult(h, rho, omega, iters, flags, RowLists(ln, self.index, off, 0, None), None, 1), 0


def _check(rc, what):
    return _hip.check(rc, what)


class CapturedEnsemble:
    """An ``ensemble.Ensemble`` whose steps replay one recorded graph.

    >>> ens = Ensemble(ics, cfg); ens.step()        # an eager step sizes the buffers
    >>> cap = CapturedEnsemble(ens)
    >>> for _ in range(100):
    ...     cap.step()
    >>> cap.sync_back()                              # ens.state is the replayed state again
    """

    def __init__(self, ens: ensemble.Ensemble, headroom: float = 1.3):
        if ens.state.row_counts is None:
            raise ValueError("take one eager step first: it sizes the recorded step's buffers")
        hc = ens.cfg.hydro
        if hc.ns_nu > 0 or hc.ns_alpha > 0 or hc.disc_alpha_av > 0:
            raise ValueError("the recorded ensemble step covers the cloud physics without disc viscosity")
        self.ens = ens
        self.cfg = ens.cfg
        self.headroom = float(headroom)
        self.replays = 0
        self.rerecords = 0  # recordings redone because library scratch moved under them
        self.fallbacks = 0
        self.fallback_reasons = []
        self._record()

    # ------------------------------------------------------------------ recording
    def _record(self):
        st = self.ens.state
        cfg = self.cfg
        n = int(st.x.shape[0])
        nseg = len(st.seg_off) - 1
        L = _hip.lib()
        sizes = np.zeros(6, np.int64)
        _check(L.nb_ctx_sizes(_hip.ctx(), sizes.ctypes.data_as(_hip._P64)), "nb_ctx_sizes")
        # what the recorded pass will ask for: row capacities from the last pass's counts, as the
        # library sizes them from a hint (a quarter + 32 entries, extras doubled + 16, multiples of 4)
        c = st.row_counts.cpu().numpy().astype(np.int64)
        a, b = c[:, 0], c[:, 1]
        need = (int(np.sum((a + a // 4 + 32 + 3) // 4 * 4)), int(np.sum((2 * b + 16 + 3) // 4 * 4)), int(np.sum(a + b)),
                int(sizes[3]))  # fmt: skip
        hr = self.headroom
        gather, extras, index, cells = (int(v * hr) + 4096 for v in need)
        self.capacities = {"gather": gather, "extras": extras, "index": index, "cells": cells}
        _check(L.nb_fixed_reserve(_hip.ctx(), n, nseg, gather, extras, index, cells, int(not (st.m_uniform > 0.0))),
               "nb_fixed_reserve")  # fmt: skip
        dev = st.x.device
        self.n, self.nseg = n, nseg
        self.seg_off_dev = torch.as_tensor(np.asarray(st.seg_off, np.int64), device=dev)
        self.status = torch.zeros(16, dtype=torch.int32, device=dev)
        self.minima = torch.empty((nseg, 4), dtype=torch.float64, device=dev)
        self.dt_seg = torch.empty(nseg, dtype=torch.float64, device=dev)
        self.index = torch.empty(index, dtype=torch.int32, device=dev)
        self.S = {k: getattr(st, k).clone() for k in _FIELDS}
        self.backup = {k: torch.empty_like(v) for k, v in self.S.items()}
        self.fixed = _Fixed(self.status, self.seg_off_dev, nseg, self.index)
        self._template = st
        # one untimed pass of the recorded sequence outside the capture (first-call set-up of the
        # library and torch), then the recording itself
        self._minima_host = self._host_minima_from(st)
        self._set_dt()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._record_body(commit=False)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph(keep_graph=True)  # kept for node_counts()
        with torch.cuda.graph(self.graph):
            self._record_body(commit=True)
        self.graph.instantiate()
        torch.cuda.synchronize()
        if self.node_counts().get("memset", 0):
            # memset nodes have misbehaved on replays after the first: refuse such a recording
            raise RuntimeError("the recorded step holds memset nodes; it is not replayed")
        self._generation = _hip.generation()

    def _record_body(self, commit: bool):
        """The recorded step: S -> S (with S saved in backup first), status and minima written."""
        S = self.S
        for k, v in S.items():
            self.backup[k].copy_(v)
        self.status.zero_()
        st = replace(self._template, **{k: S[k] for k in _FIELDS})
        new = self._step(st)
        if commit:
            for k in _FIELDS:
                S[k].copy_(getattr(new, k))
            _hip.call("nb_dt_minima_seg_dev", _hip.ptr(S["h"]), _hip.ptr(S["vsig_max"]), _hip.ptr(S["a_hydro"]), None,
                      0.25, _hip.ptr(S["x"]), _hip.ptr(self.seg_off_dev), self.nseg, _hip.ptr(self.minima))  # fmt: skip
        else:
            for k in _FIELDS:  # the untimed pass leaves the state as it was
                S[k].copy_(self.backup[k])

    def _step(self, st):
        """ensemble.step with the per-cloud dt read from self.dt_seg on the device and the
        fixed-capacity passes."""
        cfg = self.cfg
        n = self.n
        dt_p, half = _dev.empty(n), _dev.empty(n)
        _hip.call("nb_expand_seg", _hip.ptr(self.dt_seg), _hip.ptr(st.cloud), n, 1.0, _hip.ptr(dt_p))
        _hip.call("nb_expand_seg", _hip.ptr(self.dt_seg), _hip.ptr(st.cloud), n, 0.5, _hip.ptr(half))

        def kick(y, a, w):
            out = torch.empty_like(y)
            _hip.call("nb_axpy_rows", _hip.ptr(y), _hip.ptr(a), _hip.ptr(half), n, w, _hip.ptr(out))
            return out

        p_half = kick(st.p, st.a_hydro, 3)
        u_half = kick(st.u, st.dudt, 1)
        if coupled._gn(cfg):
            x1, p_flow = _dev.empty((n, 3)), _dev.empty((n, 3))
            iters = _dev.empty(n, torch.int32)
            _hip.call("nb_glf_step", _hip.ptr(st.x), _hip.ptr(p_half), _hip.ptr(dt_p), 0.0, n, _hip.ptr(x1),
                      _hip.ptr(p_flow), _hip.ptr(iters), _hip.ptr(self.status[6:]))  # fmt: skip
        else:
            x1 = _dev.empty((n, 3))
            _hip.call("nb_axpy_rows", _hip.ptr(st.x), _hip.ptr(p_half), _hip.ptr(dt_p), n, 3, _hip.ptr(x1))
            p_flow = p_half.clone()
        p_pred = kick(p_flow, st.a_hydro, 3)
        u_pred = kick(u_half, st.dudt, 1)
        order = _dev.empty(n, torch.int32)
        hs = coupled._search_lengths(cfg, st.h)
        _hip.call("nb_grid_order_seg_fixed", _hip.ptr(x1), _hip.ptr(hs), n, _hip.ptr(st.cloud),
                  _hip.ptr(self.seg_off_dev), self.nseg, float(st.cell_factor), _hip.ptr(self.status),
                  _hip.ptr(order))  # fmt: skip
        counts = st.row_counts.view(torch.int64).view(n)
        x1, p_flow, p_pred, u_half, u_pred, a_prev, mass, h, alpha, perm, half, counts = permute_many(
            (x1, p_flow, p_pred, u_half, u_pred, st.a_hydro, st.mass, st.h, st.alpha, st.perm, half, counts), order)
        pred = replace(st, x=x1, p=p_flow, u=u_pred, mass=mass, h=h, alpha=alpha, perm=perm,
                       row_counts=counts.view(torch.int32).view(n, 2))  # fmt: skip
        new = coupled._hydro(pred, cfg, coupled.velocity(cfg, x1, p_pred), 0.0, a_prev, fixed=self.fixed)
        return replace(new, p=kick(p_flow, new.a_hydro, 3), u=kick(u_half, new.dudt, 1))

    # ------------------------------------------------------------------ stepping
    @staticmethod
    def _host_minima_from(st):
        s = len(st.seg_off) - 1
        out = np.zeros(4 * s)
        _hip.call("nb_dt_minima_seg", _hip.ptr(st.h), _hip.ptr(st.vsig_max), _hip.ptr(st.a_hydro), _hip.ptr(st.nu),
                  0.25, _hip.ptr(st.x), st.seg_off.ctypes.data_as(_hip._P64), s, out.ctypes.data_as(_hip._PD))  # fmt: skip
        return out.reshape(s, 4)

    def _dt_from(self, m):
        """ensemble.stable_dt_clouds from the per-cloud minima (the same host arithmetic)."""
        hc = self.cfg.hydro
        dt = hc.courant * m[:, 0]
        dt = np.where(np.isfinite(m[:, 1]), np.minimum(dt, hc.c_force * m[:, 1]), dt)
        if self.cfg.hole:
            dt = np.minimum(dt, np.array([self.cfg.eta_orb * float(r) ** 1.5 for r in m[:, 3]]))
        return dt

    def _set_dt(self):
        self._dt = self._dt_from(self._minima_host)
        self.dt_seg.copy_(torch.as_tensor(self._dt, device=self.dt_seg.device))

    def step(self) -> np.ndarray:
        """One replayed step of every cloud; returns the steps taken."""
        dt = self._dt
        if _hip.generation() != self._generation:
            # an eager call since the recording grew (moved) library scratch the graph points at
            self.rerecords += 1
            self.sync_back()
            self._record()
            self._set_dt()
        self.graph.replay()
        packed = torch.cat([self.status.double(), self.minima.reshape(-1)]).cpu().numpy()
        status = packed[:16].astype(np.int64)
        self.replays += 1
        if np.any(status[:9] != 0):
            self._fallback(status)
        else:
            self._minima_host = packed[16:].reshape(self.nseg, 4)
        self.ens.t += dt
        self._set_dt()
        return dt

    def _fallback(self, status):
        """Restore the state saved at the replay's start, do the step eagerly, record again."""
        self.fallbacks += 1
        self.fallback_reasons.append({STATUS_NAMES[k]: int(status[k]) for k in range(9) if status[k]})
        for k, v in self.backup.items():
            self.S[k].copy_(v)
        self.sync_back()
        st = self.ens.state
        self.ens.state = ensemble.step(st, self.cfg, self._dt)
        self.headroom = max(self.headroom, 1.3)
        self._record()  # capacities from the eager step
        self._minima_host = self._host_minima_from(self.ens.state)

    def sync_back(self):
        """Put the replayed state back into the Ensemble (device copies)."""
        self.ens.state = replace(self._template, **{k: v.clone() for k, v in self.S.items()})

    def node_counts(self) -> dict:
        """Nodes of the recorded graph by type."""
        counts = np.zeros(16, np.int32)
        _check(_hip.lib().nb_graph_node_counts(ctypes_void(self.graph.raw_cuda_graph()),
                                               counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "nb_graph_node_counts")  # fmt: skip
        names = ("kernel", "memcpy", "memset", "host", "graph", "empty", "wait_event", "event_record")
        return {names[k] if k < len(names) else f"type{k}": int(c) for k, c in enumerate(counts) if c}


def ctypes_void(handle):
    return ctypes.c_void_p(int(handle))
