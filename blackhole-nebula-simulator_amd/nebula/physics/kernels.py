"""SPH smoothing kernels (physics/kernels.py of the reference): W = C_norm/h^3 f(r/h), support
2h. The GPU evaluates f and f' inside its pair loops (csrc/nb_common.h); this module holds the
kernel table, the NumPy forms used by host-side checks, and the selection by name."""

from dataclasses import dataclass

import numpy as np

R_KERN = 2.0


def cubic_f(q):
    q = np.asarray(q, float)
    return np.where(q < 1.0, 1.0 - 1.5 * q**2 + 0.75 * q**3, np.where(q < 2.0, 0.25 * (2.0 - q) ** 3, 0.0))


def cubic_df(q):
    q = np.asarray(q, float)
    return np.where(q < 1.0, -3.0 * q + 2.25 * q**2, np.where(q < 2.0, -0.75 * (2.0 - q) ** 2, 0.0))


def wendland_c4_f(q):
    q = np.asarray(q, float)
    x = np.clip(1.0 - 0.5 * q, 0.0, None)
    return np.where(q < 2.0, x**6 * (35.0 / 12.0 * q**2 + 3.0 * q + 1.0), 0.0)


def wendland_c4_df(q):
    q = np.asarray(q, float)
    x = np.clip(1.0 - 0.5 * q, 0.0, None)
    return np.where(q < 2.0, -7.0 / 3.0 * q * x**5 * (5.0 * q + 2.0), 0.0)


@dataclass(frozen=True)
class Kernel:
    name: str
    c_norm: float
    hfact: float
    f: object
    df: object

    def w(self, r, h):
        h = np.asarray(h, float)
        return self.c_norm / h**3 * self.f(np.asarray(r, float) / h)

    def dw_dr(self, r, h):
        h = np.asarray(h, float)
        return self.c_norm / h**4 * self.df(np.asarray(r, float) / h)

    def dw_dh(self, r, h):
        h = np.asarray(h, float)
        q = np.asarray(r, float) / h
        return -self.c_norm / h**4 * (3.0 * self.f(q) + q * self.df(q))

    def expected_neighbours(self, eta=None) -> float:
        eta = self.hfact if eta is None else eta
        return 4.0 * np.pi / 3.0 * (R_KERN * eta) ** 3


CUBIC = Kernel("cubic", 1.0 / np.pi, 1.2, cubic_f, cubic_df)
WENDLAND_C4 = Kernel("wendland_c4", 495.0 / (256.0 * np.pi), 1.6, wendland_c4_f, wendland_c4_df)
_KERNELS = {k.name: k for k in (CUBIC, WENDLAND_C4)}
DEFAULT_KERNEL = "wendland_c4"
DEFAULT_ETA = WENDLAND_C4.hfact
KERNEL_IDS = {"cubic": 0, "wendland_c4": 1}
C_NORM = np.array([CUBIC.c_norm, WENDLAND_C4.c_norm])


def get_kernel(name: str = DEFAULT_KERNEL) -> Kernel:
    try:
        return _KERNELS[name]
    except KeyError:
        raise ValueError(f"unknown kernel {name!r}; choose from {sorted(_KERNELS)}") from None
