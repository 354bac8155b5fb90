"""Hydro configuration and equation of state (hydro.py of the reference).

``HydroConfig`` keeps the reference's fields and defaults (hydro.py:53-98) and adds the
fields this implementation needs: ``balsara`` (Balsara limiter on the viscosity, used by the
stream-crash workload), ``precision`` ("fp64" or "mixed" pair sums) and ``learned_h`` (the
network-predicted starting h of the Newton solve).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from ..physics import units
from ..physics.kernels import DEFAULT_ETA, DEFAULT_KERNEL
from .sph import COND_PRESSURE, ShockCapturing


@dataclass(frozen=True)
class HydroConfig:
    gamma: float = 5.0 / 3.0
    isothermal_cs: float | None = None
    iso_r0: float | None = None
    iso_q: float = 0.75
    eos: str = "ideal"
    gas_units: units.GasUnits = field(default_factory=units.GasUnits)  # gas mass scale of "gasrad"
    kernel: str = DEFAULT_KERNEL
    eta: float = DEFAULT_ETA
    box: tuple | None = None
    h_max: float = np.inf
    courant: float = 0.3
    skin: float = 1.4
    c_force: float = 0.25
    viscosity: bool = True
    conductivity: bool = True
    switch: bool = True
    alpha_const: float = 1.0
    alpha_init: float = 0.0
    beta: float = 2.0
    alpha_u: float = 1.0
    conductivity_signal: int = COND_PRESSURE
    reconstruction: bool = True
    recon_closeness: bool = True
    recon_gate: int = 2
    ns_nu: float = 0.0
    ns_alpha: float = 0.0
    disc_alpha_av: float = 0.0
    # additions of this implementation
    balsara: bool = False
    precision: str = "fp64"
    learned_h: bool = False
    # the network (nebula.sim.learned_h.LearnedH) used when learned_h is set, or that collects
    # training pairs when its ``collect`` log is set
    learned_model: object = field(default=None, compare=False, repr=False)

    def shock(self) -> ShockCapturing:
        """The force pass's shock-capturing settings (hydro.py:92-98)."""
        return ShockCapturing(
            self.beta, self.alpha_u, self.conductivity_signal,
            reconstruction=self.reconstruction and self.viscosity,
            closeness=self.recon_closeness, eta_crit=1.0 / self.eta, recon_gate=self.recon_gate,
        )  # fmt: skip


def floor_energy(cfg: HydroConfig, rho, t_floor: float):
    """Code specific energy at temperature t_floor, per particle (hydro.py:155-163): one value for
    the ideal gas, a function of density for gas + radiation (nb_gasrad_u). Device tensors in give
    a device tensor out."""
    from .. import _dev, _hip

    if cfg.eos == "gasrad":
        as_dev = _dev.is_dev(rho)
        rho_d = _dev.f64(rho)
        out = torch.empty_like(rho_d)
        _hip.call("nb_gasrad_u", _hip.ptr(rho_d), float(cfg.gas_units.density_kg_m3), float(t_floor),
                  int(rho_d.numel()), _hip.ptr(out))  # fmt: skip
        return out if as_dev else _dev.host(out)
    if cfg.eos == "ideal":
        if _dev.is_dev(rho):
            return torch.full_like(rho, units.ideal_u_floor(t_floor, cfg.gamma))
        return np.full_like(np.asarray(rho, float), units.ideal_u_floor(t_floor, cfg.gamma))
    raise ValueError(f"no temperature floor for equation of state {cfg.eos!r}")


def gasrad_state_device(cfg: HydroConfig, rho_d, u_d, fixed_status=None):
    """(P, c_s, T) device tensors of the gas + radiation EOS (nb_eos_gasrad). Without
    fixed_status the failure counts are read and raised as the reference does (ValueError for u
    outside the solver bracket, RuntimeError for a solve that did not converge); with it (a
    step run without host synchronisation) they are added to its words 7 and 8."""
    from .. import _dev, _hip

    n = int(rho_d.numel())
    p, cs, t = (_dev.empty(n) for _ in range(3))
    bad = torch.zeros(2, dtype=torch.int32, device=rho_d.device) if fixed_status is None else fixed_status[7:9]
    _hip.call("nb_eos_gasrad", _hip.ptr(rho_d), _hip.ptr(u_d), float(cfg.gas_units.density_kg_m3), n, _hip.ptr(p),
              _hip.ptr(cs), _hip.ptr(t), _hip.ptr(bad))  # fmt: skip
    if fixed_status is None:
        b = bad.cpu().numpy()
        if b[0]:
            raise ValueError(f"u outside the EOS bracket [1, 1e13] K at {int(b[0])} particles")
        if b[1]:
            raise RuntimeError("gas + radiation EOS: temperature did not converge")
    return p, cs, t


def eos_state(cfg: HydroConfig, rho, u, x=None):
    """Pressure and sound speed (hydro.py:121-139) for host arrays (gas + radiation through its
    device kernel); the step itself uses the device kernels nb_eos_ideal / nb_eos_gasrad."""
    if cfg.isothermal_cs is not None:
        if cfg.iso_r0 is not None:
            cs = cfg.isothermal_cs * (np.linalg.norm(x, axis=1) / cfg.iso_r0) ** (-cfg.iso_q)
            return cs**2 * np.asarray(rho), cs
        return cfg.isothermal_cs**2 * np.asarray(rho), np.full_like(rho, cfg.isothermal_cs)
    if np.any(np.asarray(u) <= 0):
        raise FloatingPointError(f"non-positive internal energy at {int(np.sum(np.asarray(u) <= 0))} particles")
    if cfg.eos == "ideal":
        return (cfg.gamma - 1.0) * np.asarray(rho) * np.asarray(u), np.sqrt(cfg.gamma * (cfg.gamma - 1.0) * np.asarray(u))
    if cfg.eos == "gasrad":
        from .. import _dev

        p, cs, _ = gasrad_state_device(cfg, _dev.f64(rho), _dev.f64(u))
        return _dev.host(p), _dev.host(cs)
    if cfg.eos == "radiation":
        return np.asarray(rho) * np.asarray(u) / 3.0, np.sqrt(4.0 * np.asarray(u) / 9.0)
    raise ValueError(f"unsupported equation of state {cfg.eos!r}")
