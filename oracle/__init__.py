"""FP64 CPU oracle for the SPH step -- TEST INFRASTRUCTURE ONLY.

This package restates the reference simulator's hot path on the CPU so that the HIP
implementation can be checked against it, and so that bench.py can time a CPU baseline.
Only ``tests/``, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import it,
and only as the checker: the product package never imports, links or calls anything here.

Layout:
  nebula_oracle.c   the numba kernels restated in C (OpenMP, -ffp-contract=off)
  lib.py            ctypes bindings of that library
  neighbours.py     cell grid lists (neighbours.py:240-323 of the reference), brute force
  sph.py            density / forces / gradients / disc / Navier-Stokes passes (sph.py)
  integrator.py     generalized leapfrog and GN velocity (integrator.py, fastphys.py)
  physics.py        closed-form physics used by the step (eos, cooling, viscosity switch, GN H)
  coupled.py        the shared-step GN + SPH stepper (coupled.py:198-270)

Every function names the reference file:line it follows. Functions whose bitwise result
cannot be pinned to the reference (3x3 inverse and determinant, which the reference takes
from LAPACK) say so in their docstring.
"""

from . import lib  # noqa: F401  (builds nothing; fails loudly if the library is missing)
