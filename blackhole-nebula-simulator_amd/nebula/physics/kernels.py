"""Kernel selection for the HIP pair loops.

The reference's ``physics/kernels.py`` (``kernels.py:3-13, 94-102``) offers two smoothing
kernels by name, both W(r, h) = C_norm / h^3 * f(r/h) with support 2h:

* ``"cubic"``: the M4 cubic spline, C_norm = 1/pi, usual eta 1.2;
* ``"wendland_c4"``: the Wendland C4 kernel in 3D, C_norm = 495/(256 pi), usual eta 1.6 (the
  reference's default).

The GPU evaluates f and f' inside its pair loops (``csrc/nb_common.h``), selected by an integer
id; this module only maps names to those ids, their normalisations and the default eta. There
are no host-side kernel evaluations here: the FP64 oracle under ``oracle/`` has its own.
"""

from math import pi
from typing import NamedTuple

import numpy as np


class KernelSpec(NamedTuple):
    kid: int  # the id the C ABI takes (csrc/nb_common.h: 0 cubic, 1 Wendland C4)
    c_norm: float  # 3D normalisation of W = c_norm / h^3 f(q)
    eta: float  # usual smoothing-length factor h = eta (m / rho)^(1/3)


SPECS = {
    "cubic": KernelSpec(0, 1.0 / pi, 1.2),
    "wendland_c4": KernelSpec(1, 495.0 / (256.0 * pi), 1.6),
}

DEFAULT_KERNEL = "wendland_c4"  # the reference's default kernel (kernels.py:101)
DEFAULT_ETA = SPECS[DEFAULT_KERNEL].eta

# name -> id, and the normalisations indexed by id (what the wrappers pass to the library)
KERNEL_IDS = {name: s.kid for name, s in SPECS.items()}
C_NORM = np.array([s.c_norm for s in sorted(SPECS.values())])


def spec(name: str) -> KernelSpec:
    """The kernel called ``name``; ValueError for an unknown name, as the reference's lookup."""
    if name not in SPECS:
        raise ValueError(f"unknown kernel {name!r}; known kernels: {', '.join(sorted(SPECS))}")
    return SPECS[name]
