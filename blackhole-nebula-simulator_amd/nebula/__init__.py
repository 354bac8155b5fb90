"""nebula -- the SPH step of the Blackhole-Nebula-Simulator on MI355X (gfx950).

The reference's call surface for the hot path (nebula.sim.neighbours, nebula.sim.sph,
nebula.sim.integrator, nebula.sim.fastphys, nebula.sim.coupled, nebula.sim.hydro), backed by
hand-written HIP kernels in libnebula_hip.so through a ctypes C ABI (nebula._hip).
"""

__version__ = "0.1.0"
