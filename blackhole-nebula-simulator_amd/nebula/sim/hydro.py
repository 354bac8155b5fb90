"""Hydro configuration and equation of state (hydro.py of the reference).

``HydroConfig`` keeps the reference's fields and defaults (hydro.py:53-98) and adds the
fields this implementation needs: ``balsara`` (Balsara limiter on the viscosity, used by the
stream-crash workload), ``precision`` ("fp64" or "mixed" pair sums) and ``learned_h`` (the
network-predicted starting h of the Newton solve).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

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
    """Code specific energy at temperature t_floor (hydro.py:155-163, ideal branch)."""
    rho = np.asarray(rho, float)
    if cfg.eos == "ideal":
        return np.full_like(rho, units.ideal_u_floor(t_floor, cfg.gamma))
    raise ValueError(f"no temperature floor for equation of state {cfg.eos!r}")


def eos_state(cfg: HydroConfig, rho, u, x=None):
    """Pressure and sound speed (hydro.py:121-139) for host arrays; the step itself uses the
    device kernel nb_eos_ideal."""
    if cfg.isothermal_cs is not None:
        if cfg.iso_r0 is not None:
            cs = cfg.isothermal_cs * (np.linalg.norm(x, axis=1) / cfg.iso_r0) ** (-cfg.iso_q)
            return cs**2 * np.asarray(rho), cs
        return cfg.isothermal_cs**2 * np.asarray(rho), np.full_like(rho, cfg.isothermal_cs)
    if np.any(np.asarray(u) <= 0):
        raise FloatingPointError(f"non-positive internal energy at {int(np.sum(np.asarray(u) <= 0))} particles")
    if cfg.eos == "ideal":
        return (cfg.gamma - 1.0) * np.asarray(rho) * np.asarray(u), np.sqrt(cfg.gamma * (cfg.gamma - 1.0) * np.asarray(u))
    if cfg.eos == "radiation":
        return np.asarray(rho) * np.asarray(u) / 3.0, np.sqrt(4.0 * np.asarray(u) / 9.0)
    raise ValueError(f"unsupported equation of state {cfg.eos!r}")
