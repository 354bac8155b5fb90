"""Seeded initial conditions of the workloads (code units G = c = M_BH = 1).

- ``jittered_sphere_on_parabola``: a uniform sphere cut from a cubic lattice with per-axis
  uniform jitter of 0.1 spacing, the N points nearest the centre kept and rescaled to the
  radius, at r0 on the incoming branch of the parabolic (H = 0) GN orbit with pericentre r_p;
  every particle moves with the centre's velocity (as cloud.py:73 of the reference does).
- ``stream_crash``: two thin cylindrical streams of uniform density with 1% Gaussian mass
  noise, meeting at a point at radius r_meet with a given opening angle between their
  directions of motion, each at a given speed.

- ``torus``: a thick constant-angular-momentum torus in hydrostatic equilibrium in the GN
  potential, polytropic (gamma 5/3), with seeded velocity noise.

Particles are returned in the order of a cubic cell grid of about one smoothing length
(raster order, as the neighbour search traverses cells), which keeps each neighbour list
nearly sorted and the pair loops' gathers local.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .physics import units


@dataclass
class IC:
    x: np.ndarray
    v: np.ndarray
    u: np.ndarray
    mass: np.ndarray
    h: np.ndarray  # starting guess eta (m / rho)^(1/3)


def raster_order(x: np.ndarray, cell: float) -> np.ndarray:
    """Indices sorting points by the raster id of a cubic grid of the given cell size."""
    lo = x.min(axis=0)
    c = np.floor((x - lo) / cell).astype(np.int64)
    nc = c.max(axis=0) + 1
    key = (c[:, 0] * nc[1] + c[:, 1]) * nc[2] + c[:, 2]
    return np.argsort(key, kind="stable")


def gn_parabolic_state(r0: float, r_p: float) -> tuple[np.ndarray, np.ndarray]:
    """Position and velocity at radius r0 on the incoming branch of the GN orbit with energy
    H = 0 and pericentre r_p, in the x-y plane: L^2 = 2 r_p^2 / (r_p - 2) (the Schwarzschild
    marginally bound orbit), p_t = L / r0, and p_r < 0 from H = 0 with A + C r^2 = (r-2)^2/r^2."""
    ell = np.sqrt(2.0 * r_p**2 / (r_p - 2.0))
    x = np.array([r0, 0.0, 0.0])
    a = (r0 - 2.0) / r0
    p_t = ell / r0
    p_r = -np.sqrt((2.0 / r0 - a * p_t**2) / ((r0 - 2.0) ** 2 / r0**2))
    p = np.array([p_r, p_t, 0.0])
    c = -2.0 * (r0 - 2.0) / r0**4
    v = a * p + c * np.dot(p, x) * x
    return x, v


def jittered_sphere_on_parabola(n: int = 4096, radius: float = 150.0, cloud_mass_msun: float = 0.1,
                                r0: float = 20000.0, r_p: float = 2000.0, temperature: float = 1e4,
                                eta: float = 1.2, jitter: float = 0.1, seed: int = 20261014) -> IC:  # fmt: skip
    rng = np.random.default_rng(seed)
    k = int(np.ceil((n * 6.0 / np.pi) ** (1.0 / 3.0))) + 2  # lattice side whose ball holds > n
    g = np.arange(k) - (k - 1) / 2.0
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = pts + rng.uniform(-jitter, jitter, pts.shape)
    r = np.linalg.norm(pts, axis=1)
    keep = np.argsort(r, kind="stable")[:n]
    pts = pts[keep]
    pts *= radius / np.linalg.norm(pts, axis=1).max()
    m = units.msun_to_code_mass(cloud_mass_msun) / n
    rho = units.msun_to_code_mass(cloud_mass_msun) / (4.0 / 3.0 * np.pi * radius**3)
    h0 = eta * (m / rho) ** (1.0 / 3.0)
    pts = pts[raster_order(pts, 1.4 * h0)]
    xc, vc = gn_parabolic_state(r0, r_p)
    return IC(
        x=pts + xc,
        v=np.tile(vc, (n, 1)),
        u=np.full(n, units.gas_u_from_temperature(temperature)),
        mass=np.full(n, m),
        h=np.full(n, h0),
    )


def _cylinder(n: int, radius: float, length: float, rng, jitter: float = 0.1) -> np.ndarray:
    """n points filling a cylinder along +z from z = 0 to -length (jittered cubic lattice,
    the n points nearest the axis kept)."""
    vol = np.pi * radius**2 * length
    a = (vol / n) ** (1.0 / 3.0)
    for _ in range(20):
        gx = np.arange(-radius, radius + a, a)
        gz = np.arange(0.5 * a, length, a)
        X, Y, Z = np.meshgrid(gx, gx, gz, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), -Z.ravel()], axis=1)
        pts = pts + rng.uniform(-jitter * a, jitter * a, pts.shape)
        rr = np.hypot(pts[:, 0], pts[:, 1])
        if np.sum(rr < radius) >= n:
            return pts[np.argsort(rr, kind="stable")[:n]]
        a *= 0.98
    raise RuntimeError("could not fill the cylinder")


def stream_crash(n: int = 1_048_576, radius: float = 20.0, length: float = 2000.0, stream_mass_msun: float = 0.05,
                 r_meet: float = 1664.0, opening_deg: float = 160.0, speed: float = 0.056, temperature: float = 1e4,
                 noise: float = 0.01, gap: float = 2.0, eta: float = 1.2, seed: int = 20261014) -> IC:  # fmt: skip
    """Two streams of n/2 particles each. Their directions of motion d1, d2 lie in the x-y plane,
    symmetric about -x (towards the hole) with the given opening angle between them; each
    stream's leading end sits ``gap`` behind the meeting point (r_meet, 0, 0)."""
    rng = np.random.default_rng(seed)
    half = n // 2
    meet = np.array([r_meet, 0.0, 0.0])
    th = np.radians(opening_deg / 2.0)
    xs, vs, ms = [], [], []
    m_mean = units.msun_to_code_mass(stream_mass_msun) / half
    for sgn, count in ((1.0, half), (-1.0, n - half)):
        d = np.array([-np.cos(th), sgn * np.sin(th), 0.0])  # direction of motion
        body = _cylinder(count, radius, length, rng)  # along +z, extending to -length
        # rotate +z onto d: basis (e1, e2, d)
        e1 = np.array([0.0, 0.0, 1.0])
        e2 = np.cross(d, e1)
        rot = np.stack([e1, e2, d], axis=1)
        x = body @ rot.T + meet - gap * d
        xs.append(x)
        vs.append(np.tile(speed * d, (count, 1)))
        ms.append(m_mean * (1.0 + noise * rng.standard_normal(count)))
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    mass = np.concatenate(ms)
    rho = units.msun_to_code_mass(stream_mass_msun) / (np.pi * radius**2 * length)
    h0 = eta * (m_mean / rho) ** (1.0 / 3.0)
    order = raster_order(x, 1.4 * h0)
    return IC(x=x[order], v=v[order], u=np.full(n, units.gas_u_from_temperature(temperature)), mass=mass[order],
              h=np.full(n, h0))  # fmt: skip


def _torus_phi_eff(R, z, ell2):
    """Effective potential of azimuthal motion with specific angular momentum ell in the GN
    Hamiltonian: H = A |p|^2 / 2 - 1/r with p . x = 0, |p| = ell / R, A = (r - 2) / r."""
    r = np.sqrt(R * R + z * z)
    return 0.5 * (r - 2.0) / r * ell2 / (R * R) - 1.0 / r


def torus(n: int = 4_194_304, r_in: float = 12.0, r_max: float = 25.0, mass_msun: float = 0.05,
          noise: float = 0.02, gamma: float = 5.0 / 3.0, eta: float = 1.2, seed: int = 20261014) -> IC:  # fmt: skip
    """Thick torus of constant specific angular momentum ell, ell^2 = r_max^2 / (r_max - 3) (the
    circular orbit at r_max, where the effective potential of the GN Hamiltonian has its
    minimum in the plane), inner edge r_in. Hydrostatic polytrope P = K rho^gamma: the enthalpy
    w = gamma / (gamma - 1) K rho^(gamma - 1) = Phi_eff(r_in, 0) - Phi_eff(R, z) inside, so the
    specific energy u = w / gamma follows from the geometry alone (K scales the density only;
    K is fixed by the torus mass). Equal-mass particles sampled with probability proportional
    to rho (seeded rejection sampling in the bounding cylinder); velocity = dH/dp of
    p = ell / R along phi, plus Gaussian noise of ``noise`` times its size per component;
    starting h from the distance to the particle's 58th neighbour."""
    rng = np.random.default_rng(seed)
    ell2 = r_max * r_max / (r_max - 3.0)
    phi_in = _torus_phi_eff(r_in, 0.0, ell2)
    w_max = phi_in - _torus_phi_eff(r_max, 0.0, ell2)
    # outer edge in the plane and half height at r_max (w = 0)
    Rg = np.linspace(r_max, 50.0 * r_max, 200001)
    r_out = float(Rg[np.argmax(phi_in - _torus_phi_eff(Rg, 0.0, ell2) <= 0.0)])
    zg = np.linspace(0.0, r_out, 200001)
    z_top = float(zg[np.argmax(phi_in - _torus_phi_eff(r_max, zg, ell2) <= 0.0)])
    z_top = max(z_top, float(np.max(np.where(phi_in - _torus_phi_eff(np.linspace(r_in, r_out, 401)[:, None],
                                                                         zg[None, ::500], ell2) > 0,
                                                zg[None, ::500], 0.0))))  # fmt: skip
    xs = []
    got = 0
    while got < n:
        m = 4 * (n - got) + 1024
        R = np.sqrt(rng.uniform(r_in * r_in, r_out * r_out, m))  # uniform in volume
        z = rng.uniform(-z_top, z_top, m)
        w = phi_in - _torus_phi_eff(R, z, ell2)
        keep = (w > 0.0) & (rng.uniform(0.0, 1.0, m) < (np.maximum(w, 0.0) / w_max) ** (1.0 / (gamma - 1.0)))
        ph = rng.uniform(0.0, 2.0 * np.pi, int(keep.sum()))
        xs.append(np.stack([R[keep] * np.cos(ph), R[keep] * np.sin(ph), z[keep]], axis=1))
        got += int(keep.sum())
    x = np.concatenate(xs)[:n]
    R = np.hypot(x[:, 0], x[:, 1])
    r = np.linalg.norm(x, axis=1)
    w = phi_in - _torus_phi_eff(R, x[:, 2], ell2)
    shape = (w / w_max) ** (1.0 / (gamma - 1.0))  # rho / rho_max
    # rho_max from the mass: M = rho_max * integral of shape dV (Monte Carlo over the cylinder)
    mc = 4_000_000
    Rm = np.sqrt(rng.uniform(r_in * r_in, r_out * r_out, mc))
    zm = rng.uniform(-z_top, z_top, mc)
    wm = np.maximum(phi_in - _torus_phi_eff(Rm, zm, ell2), 0.0)
    vol = np.pi * (r_out * r_out - r_in * r_in) * 2.0 * z_top
    integral = vol * float(np.mean((wm / w_max) ** (1.0 / (gamma - 1.0))))
    total = units.msun_to_code_mass(mass_msun)
    rho = total / integral * shape
    m = total / n
    # p along phi with |p| = ell / R; v = A p (p . x = 0), A = (r - 2) / r
    ell = np.sqrt(ell2)
    phat = np.stack([-x[:, 1] / R, x[:, 0] / R, np.zeros(n)], axis=1)
    v = ((r - 2.0) / r * ell / R)[:, None] * phat
    v = v + noise * np.linalg.norm(v, axis=1, keepdims=True) * rng.standard_normal((n, 3))
    u = w / gamma
    # h from the sampled particles themselves (half the distance to the 58th neighbour: the
    # cubic kernel's support holds about 58 neighbours at eta 1.2), not from the analytic
    # density: random sampling makes the local density differ from it, and a start far from
    # the consistent h can send the first Newton step of the density solve towards h = 0
    from scipy.spatial import cKDTree

    k = int(round(4.0 / 3.0 * np.pi * (2.0 * eta) ** 3))
    dist, _ = cKDTree(x).query(x, k=[min(k, n)], workers=-1)
    h = np.minimum(0.5 * dist[:, 0], eta * (m / np.maximum(rho, 1e-300)) ** (1.0 / 3.0) * 4.0)
    order = raster_order(x, 1.4 * float(np.median(h)))
    return IC(x=x[order], v=v[order], u=u[order], mass=np.full(n, m), h=h[order])
