"""Learned starting h for the h-rho Newton solve.

A small MLP (8 -> 64 -> 64 -> 1, tanh) predicts delta = log h_final - log h_prev of every particle
from the previous step's state:

    log h, log rho, div v dt, |curl v| dt, log(neighbour count), log r, log u, dt / dt_prev.

With ``residual`` (the default) the network learns the correction to the continuity-equation
predictor d log h = div v dt / 3 (h rho^(1/3) fixed, d log rho / dt = -div v): delta = div v dt / 3
+ MLP(features). An untrained or zero network then still starts from the physical predictor.

The Newton iteration of the density pass then starts from h_prev exp(delta) instead of h_prev.
It still runs to the same tolerance and stays the arbiter; a prediction that is not finite, or
that changes h by a factor 2 or more, falls back to h_prev (counted). The search lists are
built from h_prev as before, so the lists are the plain method's.

Inference is the BF16 MFMA kernel ``nb_mlp_infer``. Training uses this project's own FP32
kernels (``nb_mlp_train_step``: forward, backward and Adam). The training pairs are logged by
the simulator itself (``PairLog``, filled by ``coupled.step`` when ``collect`` is set). With
``online`` set, the model keeps learning while it is used: ``coupled.step`` hands it every
step's pairs (``observe``), which go into a ring buffer of the latest pairs, and every few steps
a few Adam steps run on it, so the network follows the flow as it changes.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _hip

N_FEATURES = 8


def param_count() -> int:
    return int(_hip.lib().nb_mlp_param_count())


class PairLog:
    """(features, target) pairs logged from simulation steps (device tensors, float32)."""

    def __init__(self, stride: int = 1):
        self.stride = int(stride)
        self._feat: list[torch.Tensor] = []
        self._target: list[torch.Tensor] = []

    def add(self, feat: torch.Tensor, target: torch.Tensor) -> None:
        # copies: a recorded step hands in buffers of its graph, which every replay overwrites
        self._feat.append(feat.clone())
        self._target.append(target.clone())

    @property
    def size(self) -> int:
        return int(sum(t.numel() for t in self._target))

    def tensors(self) -> tuple[torch.Tensor, torch.Tensor]:
        return torch.cat(self._feat), torch.cat(self._target)


class OnlineSchedule:
    """Online training while the model is used: the last ``capacity`` pairs are kept (pairs of
    every ``stride``-th particle); every ``every`` steps, ``steps`` Adam steps of ``batch`` pairs
    at learning rate ``lr``."""

    def __init__(self, stride: int = 1, capacity: int = 131072, every: int = 4, steps: int = 8, batch: int = 4096,
                 lr: float = 1e-3):  # fmt: skip
        self.stride, self.capacity, self.every = int(stride), int(capacity), int(every)
        self.steps, self.batch, self.lr = int(steps), int(batch), float(lr)


class StartGate:
    """When the learned start runs: only while plain Newton needs it.

    The network and its online training cost a fixed time per step (inference, features,
    training). They pay only when many rows would otherwise need more than one Newton step; where
    the plain start (h_prev) already converges in one step for nearly every row, they cost time and
    save nothing. The gate watches the fraction of rows left for the slow solve (more than one
    Newton step) under the plain start:

    - the first step, and every ``probe_every``-th step while the network runs, start from h_prev
      (a probe) and measure that fraction;
    - the network runs while the probed fraction is at least ``off_below``, and is switched on
      again once a plain step's fraction exceeds ``on_above``.

    Decisions depend only on the run's own iteration counts, so equal runs decide alike. A step
    that starts from h_prev is exactly the plain method's step. A decision uses the counts of every
    step but the last one: the last step may still be running on the device when the next one is
    launched (a pipelined driver), and reading its count would wait for it.
    """

    def __init__(self, on_above: float = 0.05, off_below: float = 0.03, probe_every: int = 32):
        self.on_above, self.off_below, self.probe_every = float(on_above), float(off_below), int(probe_every)
        self.on = None  # unknown until the first probe
        self.since_probe = 0
        self._pending = []  # (device count of slow rows, rows, network used) of the steps not yet read
        self.last_plain_fraction = None
        self.steps_network = 0
        self.steps_plain = 0
        self.probes = 0

    def observe(self, slow_rows, rows: int, used_network: bool) -> None:
        """A step's count of rows left for the slow solve: a number, or a device scalar, copied to
        pinned host memory behind the step's work and read only when a later decision needs it."""
        if isinstance(slow_rows, torch.Tensor) and slow_rows.is_cuda:
            host = torch.empty((), dtype=slow_rows.dtype).pin_memory()
            host.copy_(slow_rows, non_blocking=True)
            ready = torch.cuda.Event()
            ready.record()
            slow_rows = (host, ready)
        self._pending.append((slow_rows, int(rows), bool(used_network)))

    def _resolve(self) -> None:
        while len(self._pending) > 1:  # all but the newest count
            count, rows, used = self._pending.pop(0)
            if used:
                continue
            if isinstance(count, tuple):  # (pinned copy, the event after it): waits for that step only
                count[1].synchronize()
                count = count[0]
            frac = float(count.item() if hasattr(count, "item") else count) / max(rows, 1)
            self.last_plain_fraction = frac
            if self.on:  # a probe while the network runs
                self.on = frac >= self.off_below
            else:
                self.on = frac > self.on_above

    def snapshot(self) -> dict:
        """The gate's state, for restore() when a step is done again."""
        d = dict(self.__dict__)
        d["_pending"] = list(self._pending)
        return d

    def restore(self, snap: dict) -> None:
        self.__dict__.update(snap)
        self._pending = list(snap["_pending"])

    def decide(self) -> bool:
        """True: the coming step starts from the network's prediction; False: from h_prev."""
        self._resolve()
        use = False
        if self.on is None:
            self.probes += 1
        elif self.on:
            if self.since_probe >= self.probe_every:
                self.since_probe = 0
                self.probes += 1
            else:
                self.since_probe += 1
                use = True
        if use:
            self.steps_network += 1
        else:
            self.steps_plain += 1
        return use

    def stats(self) -> dict:
        return {"steps_network": self.steps_network, "steps_plain": self.steps_plain, "probes": self.probes,
                "last_plain_slow_fraction": self.last_plain_fraction, "on": self.on}  # fmt: skip


class LearnedH:
    """The network, its normalisation and its Adam state (device tensors). ``gate`` decides per
    step whether the network runs (StartGate); None runs it on every step."""

    def __init__(self, seed: int = 0, residual: bool = True, gate: StartGate | None = None):
        self.residual = bool(residual)
        n = param_count()
        rng = np.random.default_rng(seed)
        p = np.zeros(n, np.float32)
        # W1 (8 x 64), W2 (64 x 64), w3 (64): Glorot-uniform; biases zero
        p[0:512] = rng.uniform(-1, 1, 512) * np.sqrt(6.0 / (8 + 64))
        p[576:4672] = rng.uniform(-1, 1, 4096) * np.sqrt(6.0 / (64 + 64))
        p[4736:4800] = rng.uniform(-1, 1, 64) * np.sqrt(6.0 / (64 + 1))
        dev = _hip.torch_device()
        self.params = torch.as_tensor(p, device=dev)
        self.norm = torch.as_tensor(np.concatenate([np.zeros(8), np.ones(8), [0.0, 1.0]]).astype(np.float32), device=dev)
        self.adam_m = torch.zeros(n, dtype=torch.float32, device=dev)
        self.adam_v = torch.zeros(n, dtype=torch.float32, device=dev)
        self.step_count = 0
        self.collect: PairLog | None = None
        self.last_fallbacks = None  # device count of the last h_start call
        self.online: OnlineSchedule | None = None
        self._ring = None
        self.gate = gate

    def start_online(self, **kw) -> None:
        """Keep training while the model is used (OnlineSchedule fields as keywords)."""
        self.online = OnlineSchedule(**kw)
        cap = self.online.capacity
        self._ring = [_dev.empty((cap, N_FEATURES), torch.float32), _dev.empty(cap, torch.float32), 0, 0, 0]

    def observe(self, feat: torch.Tensor, target: torch.Tensor) -> None:
        """One step's (features, converged delta) pairs: into the ring buffer; every
        ``online.every`` calls, ``online.steps`` Adam steps on random batches of it."""
        rf, rt, pos, size, calls = self._ring
        m = int(target.numel())
        cap = int(rt.numel())
        if m > cap:
            feat, target, m = feat[:cap], target[:cap], cap
        a = min(m, cap - pos)
        rf[pos:pos + a] = feat[:a]
        rt[pos:pos + a] = target[:a]
        if a < m:
            rf[: m - a] = feat[a:m]
            rt[: m - a] = target[a:m]
        pos = (pos + m) % cap
        size = min(size + m, cap)
        calls += 1
        self._ring = [rf, rt, pos, size, calls]
        o = self.online
        if calls % o.every == 0:
            self.train(rf[:size], rt[:size], steps=o.steps, batch=min(o.batch, size), lr=o.lr, seed=calls,
                       record=False)  # fmt: skip

    # ------------------------------------------------------------------ data
    def features(self, h_prev, rho_prev, div_v, curl_v, nneigh, x, u, dt, dt_prev, stride: int = 1,
                 dts=None) -> torch.Tensor:  # fmt: skip
        """Raw features of every stride-th particle. dts: (dt, dt_prev) in device memory (a step
        recorded for replay), replacing the host values."""
        n = int(h_prev.shape[0])
        rows = (n + stride - 1) // stride
        out = _dev.empty((rows, N_FEATURES), torch.float32)
        if dts is None:
            _hip.call("nb_mlp_features", _hip.ptr(h_prev), _hip.ptr(rho_prev), _hip.ptr(div_v), _hip.ptr(curl_v),
                      _hip.ptr(nneigh), _hip.ptr(x), _hip.ptr(u), n, int(stride), float(dt), float(dt_prev),
                      _hip.ptr(out))  # fmt: skip
        else:
            _hip.call("nb_mlp_features_dev", _hip.ptr(h_prev), _hip.ptr(rho_prev), _hip.ptr(div_v), _hip.ptr(curl_v),
                      _hip.ptr(nneigh), _hip.ptr(x), _hip.ptr(u), n, int(stride), _hip.ptr(dts),
                      _hip.ptr(out))  # fmt: skip
        return out

    @staticmethod
    def targets(h_prev, h_final, stride: int = 1) -> torch.Tensor:
        n = int(h_prev.shape[0])
        out = _dev.empty((n + stride - 1) // stride, torch.float32)
        _hip.call("nb_mlp_target", _hip.ptr(h_prev), _hip.ptr(h_final), n, int(stride), _hip.ptr(out))
        return out

    def _base(self, feat: torch.Tensor) -> torch.Tensor | None:
        """The continuity-equation predictor div v dt / 3 of each row (None without ``residual``)."""
        return feat[:, 2] * (1.0 / 3.0) if self.residual else None

    def net_target(self, feat: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
        """What the network itself is trained on: target minus the predictor."""
        b = self._base(feat)
        return target if b is None else (target - b).contiguous()

    # ------------------------------------------------------------------ inference
    def predict(self, feat: torch.Tensor) -> torch.Tensor:
        """delta = log h_final - log h_prev for each feature row (network: BF16 MFMA)."""
        n = int(feat.shape[0])
        out = _dev.empty(n, torch.float32)
        _hip.call("nb_mlp_infer", _hip.ptr(self.params), _hip.ptr(self.norm), _hip.ptr(feat), n, _hip.ptr(out))
        b = self._base(feat)
        return out if b is None else out + b

    def h_start(self, h_prev: torch.Tensor, feat: torch.Tensor, gate: float = 1e-4) -> torch.Tensor:
        """Starting h of the Newton solve: h_prev exp(delta), or h_prev where the predicted change
        is below ``gate`` (the Newton tolerance: one step from h_prev suffices there).
        ``last_fallbacks`` (a device int64 tensor, read lazily) counts the particles whose
        prediction was rejected (not finite, or a change by a factor 2 or more)."""
        delta = self.predict(feat)
        out = torch.empty_like(h_prev)
        fb = _dev.empty(1, torch.int64)
        _hip.call("nb_mlp_h_start", _hip.ptr(h_prev), _hip.ptr(delta), int(h_prev.shape[0]), float(gate),
                  _hip.ptr(out), _hip.ptr(fb))  # fmt: skip
        self.last_fallbacks = fb
        return out

    # ------------------------------------------------------------------ training
    def fit_normalisation(self, feat: torch.Tensor, target: torch.Tensor) -> None:
        """Feature means and 1/std, (network) target mean and std, from a training set."""
        f = feat.double()
        mean = f.mean(0)
        std = f.std(0).clamp_min(1e-12)
        t = self.net_target(feat, target).double()
        tstd = float(t.std().clamp_min(1e-12)) if t.numel() > 1 else 1.0
        vals = torch.cat([mean, 1.0 / std, torch.tensor([float(t.mean()), tstd], dtype=torch.float64, device=f.device)])
        self.norm = vals.float().contiguous()

    def grad(self, feat: torch.Tensor, target: torch.Tensor) -> tuple[float, torch.Tensor]:
        """Loss and FP32 gradient over the batch (no update)."""
        target = self.net_target(feat, target)
        g = torch.empty_like(self.params)
        loss = _dev.empty(1, torch.float32)
        feat, target = feat.contiguous(), target.contiguous()  # referenced until the call returns
        _hip.call("nb_mlp_grad", _hip.ptr(self.params), _hip.ptr(self.norm), _hip.ptr(feat), _hip.ptr(target),
                  int(target.numel()), _hip.ptr(g), _hip.ptr(loss))  # fmt: skip
        return float(loss.item()), g

    def train(self, feat: torch.Tensor, target: torch.Tensor, steps: int = 400, batch: int = 4096,
              lr: float = 3e-3, seed: int = 0, record: bool = True) -> list[float]:  # fmt: skip
        """Adam on random minibatches; returns the loss of each step (normalised target), or
        nothing without ``record`` (no host synchronisation)."""
        target = self.net_target(feat, target)
        gen = torch.Generator(device=feat.device)
        gen.manual_seed(seed)
        n = int(target.numel())
        batch = min(batch, n)
        losses = []
        loss = _dev.empty(1, torch.float32)
        for _ in range(steps):
            idx = torch.randint(0, n, (batch,), generator=gen, device=feat.device)
            fb = feat.index_select(0, idx).contiguous()
            tb = target.index_select(0, idx).contiguous()
            self.step_count += 1
            _hip.call("nb_mlp_train_step", _hip.ptr(self.params), _hip.ptr(self.adam_m), _hip.ptr(self.adam_v),
                      self.step_count, _hip.ptr(self.norm), _hip.ptr(fb), _hip.ptr(tb), batch, float(lr),
                      _hip.ptr(loss))  # fmt: skip
            if record:
                losses.append(float(loss.item()))
        return losses

    # ------------------------------------------------------------------ persistence
    def save(self, path) -> None:
        np.savez(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy(),
                 residual=np.array(self.residual))  # fmt: skip

    @classmethod
    def load(cls, path) -> "LearnedH":
        d = np.load(path, allow_pickle=False)
        m = cls(residual=bool(d["residual"]) if "residual" in d else True)
        m.params = torch.as_tensor(d["params"], device=m.params.device).contiguous()
        m.norm = torch.as_tensor(d["norm"], device=m.params.device).contiguous()
        return m
