"""Pinning the FP64 CPU oracle (CPU only): kernel closed forms and expected values, list
definitions, an independent NumPy evaluation of the density and force formulas (the reference's
readable twins, sph.py:158-209 and 444-516), lattice invariants and conservation laws. Each check
is chosen so that a plausible restatement mistake (a dropped term, a wrong sign or index, a
transposed operand) fails it."""

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.spatial import cKDTree

from helpers import check

from oracle import neighbours as on
from oracle import physics as oph
from oracle import sph as osph

CUBIC_C = 1.0 / np.pi
WEND_C = 495.0 / (256.0 * np.pi)


def fcc_lattice(n_cells, box=1.0):
    """ic.py:19-27 of the reference."""
    a = box / n_cells
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    c = np.arange(n_cells)
    cells = np.stack(np.meshgrid(c, c, c, indexing="ij"), axis=-1).reshape(-1, 3)
    return (cells[:, None, :] + basis[None, :, :] + 0.25).reshape(-1, 3) * a


def blob(rng, n=2000):
    core = rng.normal(0.0, 0.15, size=(n // 2, 3))
    bg = rng.uniform(-1.0, 1.0, size=(n - n // 2, 3))
    return np.concatenate([core, bg]), np.full(n, 1.0 / n)


# ----------------------------------------------------------------- kernels
def test_kernel_expected_values():
    """expected-values.md §4: W(0) h^3 = 0.3183098862, W(h) h^3 = 0.0795774715 (cubic)."""
    check("cubic W(0) h^3", 0.3183098862, CUBIC_C * osph.kernel_f(0.0, 0), rtol=1e-10)
    check("cubic W(h) h^3", 0.0795774715, CUBIC_C * osph.kernel_f(1.0, 0), rtol=1e-9)
    check("expected neighbours eta 1.2: (4 pi/3) 2.4^3", 57.9058, 4 * np.pi / 3 * 2.4**3, rtol=1e-5)


@pytest.mark.parametrize("kid,cn", [(0, CUBIC_C), (1, WEND_C)])
def test_kernel_normalised_and_derivative(kid, cn):
    norm = quad(lambda q: 4 * np.pi * q * q * cn * osph.kernel_f(q, kid), 0, 2, points=[1.0], epsabs=1e-13)[0]
    check(f"kid {kid}: integral of W = 1", 1.0, norm, rtol=1e-10)
    q = np.linspace(0.01, 1.99, 397)
    e = 1e-6
    fd = np.array([(osph.kernel_f(x + e, kid) - osph.kernel_f(x - e, kid)) / (2 * e) for x in q])
    an = np.array([osph.kernel_df(x, kid) for x in q])
    check(f"kid {kid}: f' = finite difference of f", fd, an, atol=1e-8)
    check(f"kid {kid}: zero beyond 2h", 0.0, osph.kernel_f(2.0, kid) + osph.kernel_df(2.5, kid))


# ----------------------------------------------------------------- lists
def _uneven(rng, n, box):
    pos = rng.uniform(0, box, size=(n, 3))
    h = np.exp(rng.uniform(np.log(0.01), np.log(0.3), n)) * box / 8
    return pos, h


@pytest.mark.parametrize("periodic", [False, True])
def test_lists_equal_brute_force_symmetric(rng, periodic):
    box = 2.0
    pos, h = _uneven(rng, 1500, box)
    b = box if periodic else None
    nb = on.find_neighbours_grid(pos, h, box=b)
    ref = on.brute_force_neighbours(pos, h, box=b)
    check("lists = brute force", True, all(np.array_equal(nb.of(i), ref[i]) for i in range(len(pos))))
    pairs = {(i, int(j)) for i in range(len(pos)) for j in nb.of(i)}
    check("symmetric", True, all((j, i) in pairs for i, j in pairs))
    check("self listed", True, all(i in set(nb.of(i)) for i in range(len(pos))))
    scatter = sum(int(np.sum(np.linalg.norm(on.minimum_image(pos[nb.of(i)] - pos[i], nb.box), axis=1) >= 2 * h[i]))
                  for i in range(len(pos)))  # fmt: skip
    check("scatter partners present", True, scatter > 0)


def test_lists_equal_kdtree(rng):
    """scipy's cKDTree (the reference's own twin, neighbours.py:53-71) as an independent search."""
    pos, h = _uneven(rng, 3000, 2.0)
    nb = on.find_neighbours_grid(pos, h)
    tree = cKDTree(pos)
    lists = tree.query_ball_point(pos, r=2.0 * h, return_sorted=False)
    rows = np.repeat(np.arange(len(pos)), [len(x) for x in lists])
    cols = np.concatenate([np.asarray(x, np.int64) for x in lists])
    import scipy.sparse as sp

    g = sp.csr_matrix((np.ones(len(cols), np.int8), (rows, cols)), shape=(len(pos), len(pos)))
    sym = (g + g.T).tocsr()
    sym.sort_indices()
    check("grid lists = cKDTree lists (offsets)", True, bool(np.array_equal(sym.indptr, nb.offset)))
    check("grid lists = cKDTree lists (indices)", True, bool(np.array_equal(sym.indices, nb.index)))


def test_lattice_neighbour_count():
    pos = fcc_lattice(8)
    h = np.full(len(pos), 1.2 * (1.0 / len(pos)) ** (1 / 3))
    counts = on.find_neighbours_grid(pos, h, box=1.0).counts - 1
    check("fcc eta 1.2: neighbours vs 57.9", 57.9, counts.mean(), rtol=0.1)
    check("all counts equal", 0, int(np.ptp(counts)))


def test_tighten_equals_direct_build(rng):
    pos, h = _uneven(rng, 3000, 2.0)
    tight = on.tighten(on.find_neighbours_grid(pos, 1.25 * h, box=2.0), pos, h)
    direct = on.find_neighbours_grid(pos, h, box=2.0)
    check("tightened = built at h", True, bool(np.array_equal(tight.offset, direct.offset)
                                                and np.array_equal(tight.index, direct.index)))  # fmt: skip


def test_periodic_box_too_small():
    with pytest.raises(ValueError, match="half the periodic box"):
        on.find_neighbours_grid(np.zeros((2, 3)) + 0.5, np.array([0.3, 0.3]), box=1.0)


# ----------------------------------------------------------------- density
def density_reference(pos, mass, h_start, nb, h_search, kid, cn, eta, tol, h_max):
    """sph.py:158-209 (the reference's readable twin), NumPy per particle."""
    def f(q):
        return np.array([osph.kernel_f(x, kid) for x in q])

    def df(q):
        return np.array([osph.kernel_df(x, kid) for x in q])

    def sums(r, m, h):
        q = r / h
        return np.sum(m * cn / h**3 * f(q)), np.sum(m * -cn / h**4 * (3.0 * f(q) + q * df(q)))

    n = len(pos)
    out = np.empty(n), np.empty(n), np.empty(n), np.zeros(n, np.int64), np.zeros(n, np.int64)
    for i in range(n):
        j = nb.of(i)
        r = np.linalg.norm(pos[j] - pos[i], axis=1)
        m_j, m_i = mass[j], mass[i]
        hi, lo = min(h_search[i], h_max), 0.0
        h = min(h_start[i], hi)
        h0 = h
        flag = 3
        for it in range(1, 51):
            rs, dws = sums(r, m_j, h)
            fv = rs - m_i * (eta / h) ** 3
            dfv = dws + 3.0 * m_i * eta**3 / h**4
            if fv > 0:
                hi = h
            else:
                lo = h
            h_new = h - fv / dfv if dfv > 0 else 0.5 * (lo + hi)
            if not lo <= h_new <= hi:
                h_new = 0.5 * (lo + hi)
            out[3][i] = it
            if abs(h_new - h) < tol * h0:
                h, flag = h_new, 0
                break
            h = h_new
        h_lim = min(h_search[i], h_max)
        if h >= h_lim * (1 - 1e-3) and sums(r, m_j, h_lim)[0] - m_i * (eta / h_lim) ** 3 < 0:
            h, flag = h_lim, (1 if h_lim >= h_max else 2)
        rs, dws = sums(r, m_j, h)
        out[0][i], out[1][i], out[4][i] = h, rs, flag
        out[2][i] = 1.0 if flag == 1 else 1.0 + h / (3.0 * rs) * dws
    return out


def test_density_matches_readable_twin(rng):
    pos, mass = blob(rng, 600)
    first = osph.compute_density(pos, mass, np.full(len(pos), 0.08), kernel_name="cubic", eta=1.2, h_max=0.5)
    hs = first.h * 1.25
    nb = on.find_neighbours_grid(pos, hs)
    res = osph.density_numba(pos, mass, first.h, nb, hs, "cubic", 1.2, 1e-12, 0.5)
    h, rho, om, _, fl = density_reference(pos, mass, first.h, nb, hs, 0, CUBIC_C, 1.2, 1e-12, 0.5)
    check("twin: h", h, res.h, rtol=1e-12)
    check("twin: rho", rho, res.rho, rtol=1e-12)
    check("twin: omega", om, res.omega, rtol=1e-12)
    check("twin: flags", fl, res.flags)


@pytest.mark.parametrize("kernel,eta", [("cubic", 1.2), ("wendland_c4", 1.6)])
def test_lattice_density(kernel, eta):
    pos = fcc_lattice(10)
    n = len(pos)
    h0 = np.full(n, eta * (1.0 / n) ** (1 / 3) * 1.1)
    res = osph.compute_density(pos, np.full(n, 1.0 / n), h0, box=1.0, kernel_name=kernel, eta=eta)
    check(f"{kernel}: lattice rho = 1 (0.5%)", np.ones(n), res.rho, rtol=5e-3)
    check(f"{kernel}: all converged", n, int(np.sum(res.flags == 0)))
    check(f"{kernel}: h = eta (m/rho)^(1/3)", eta * (1.0 / n / res.rho) ** (1 / 3), res.h, rtol=1e-4)


def test_omega_matches_finite_difference(rng):
    pos, mass = blob(rng)
    res = osph.compute_density(pos, mass, np.full(len(pos), 0.05), kernel_name="cubic", eta=1.2, h_max=0.5,
                               tol=1e-12)  # fmt: skip
    nb = on.find_neighbours_grid(pos, np.full(len(pos), 0.6))
    fd = []
    idx = list(range(0, 40)) + list(range(1000, 1040))
    for i in idx:
        r = np.linalg.norm(pos[nb.of(i)] - pos[i], axis=1)
        h, e = res.h[i], 1e-6 * res.h[i]

        def rs(hh):
            return np.sum(mass[nb.of(i)] * CUBIC_C / hh**3 * np.array([osph.kernel_f(x / hh, 0) for x in r]))

        fd.append(1 + h / (3 * res.rho[i]) * (rs(h + e) - rs(h - e)) / (2 * e))
    check("Omega vs central differences", np.array(fd), res.omega[idx], rtol=1e-6)


def test_isolated_particle_capped():
    pos = fcc_lattice(8)
    pos = np.concatenate([pos[pos[:, 0] < 0.5], [[5.0, 5.0, 5.0]]])
    n = len(pos)
    res = osph.compute_density(pos, np.full(n, 1 / 2048), np.full(n, 1.2 / 8 / 4 ** (1 / 3)), kernel_name="cubic",
                               eta=1.2, h_max=0.4)  # fmt: skip
    check("edges converge", n - 1, int(np.sum(res.flags[:-1] == 0)))
    check("isolated: flag AT_H_MAX", 1, int(res.flags[-1]))
    check("isolated: h = h_max", 0.4, res.h[-1], rtol=1e-12)
    check("isolated: Omega = 1", 1.0, res.omega[-1])


# ----------------------------------------------------------------- forces
def forces_reference(pos, vel, mass, h, rho, omega, pressure, nb, u, c_s, alpha, beta, alpha_u, sig):
    """sph.py:444-516 (the reference's readable twin, no reconstruction), NumPy per particle."""
    def dw(r, hh):
        return CUBIC_C / hh**4 * np.array([osph.kernel_df(x, 0) for x in r / hh])

    n = len(pos)
    acc, dudt, vs, vh, ch = np.zeros((n, 3)), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    for a in range(n):
        b = nb.of(a)
        b = b[b != a]
        dx = pos[a] - pos[b]
        r = np.linalg.norm(dx, axis=1)
        rhat = dx / r[:, None]
        fa, fb = dw(r, h[a]), dw(r, h[b])
        vdotr = np.sum((vel[a] - vel[b]) * rhat, axis=1)
        num_a, num_b = np.full(len(b), pressure[a]), pressure[b].astype(float)
        heat = pressure[a] / (omega[a] * rho[a] ** 2) * np.sum(mass[b] * vdotr * fa)
        on_ = vdotr < 0
        vsa = alpha[a] * c_s[a] + beta * np.abs(vdotr)
        vsb = alpha[b] * c_s[b] + beta * np.abs(vdotr)
        num_a = num_a + np.where(on_, -0.5 * rho[a] * vsa * vdotr, 0.0)
        num_b = num_b + np.where(on_, -0.5 * rho[b] * vsb * vdotr, 0.0)
        visc = -np.sum((mass[b] * vsa * 0.5 * vdotr * vdotr * fa)[on_]) / (omega[a] * rho[a])
        heat += visc
        vh[a] = visc
        vs[a] = max(np.max(vsa), c_s[a]) if len(b) else c_s[a]
        vu = np.sqrt(np.abs(pressure[a] - pressure[b]) / (0.5 * (rho[a] + rho[b]))) if sig == 0 else np.abs(vdotr)
        cond = np.sum(mass[b] * alpha_u * vu * (u[a] - u[b]) * 0.5 * (fa / (omega[a] * rho[a]) + fb / (omega[b] * rho[b])))
        heat += cond
        ch[a] = cond
        pair = mass[b] * (num_a / (omega[a] * rho[a] ** 2) * fa + num_b / (omega[b] * rho[b] ** 2) * fb)
        acc[a] = -np.sum(pair[:, None] * rhat, axis=0)
        dudt[a] = heat
    return acc, dudt, vs, vh, ch


def _moving(rng, n=800):
    pos, mass = blob(rng, n)
    d = osph.compute_density(pos, mass, np.full(n, 0.08), kernel_name="cubic", eta=1.2, h_max=0.5)
    u = rng.uniform(0.5, 1.5, n)
    vel = rng.normal(0.0, 0.3, size=pos.shape)
    p = (2.0 / 3.0) * d.rho * u
    cs = np.sqrt(10.0 / 9.0 * u)
    alpha = rng.uniform(0, 1, n)
    return pos, vel, mass, d, p, u, cs, alpha, on.tighten(d.neighbours, pos, d.h)


@pytest.mark.parametrize("sig", [0, 1])
def test_forces_match_readable_twin(rng, sig):
    pos, vel, mass, d, p, u, cs, alpha, nb = _moving(rng)
    f = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", u=u, c_s=cs, alpha=alpha,
                          alpha_u=0.05, cond_signal=sig)  # fmt: skip
    acc, dudt, vs, vh, ch = forces_reference(pos, vel, mass, d.h, d.rho, d.omega, p, nb, u, cs, alpha, 2.0, 0.05, sig)
    check("twin: acc (per vector)", acc, f.acc, rtol=1e-12, per_vector=True)
    s = np.abs(dudt).max()
    check("twin: du/dt (/max)", dudt / s, f.dudt / s, atol=1e-12)
    check("twin: v_sig max", vs, f.vsig_max, rtol=1e-14)
    check("twin: viscous heating (/max)", vh / s, f.dudt_visc / s, atol=1e-12)
    check("twin: conduction (/max)", ch / s, f.dudt_cond / s, atol=1e-12)


def test_lattice_zero_force():
    pos = fcc_lattice(8)
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 1.2 * (1.0 / n) ** (1 / 3)), box=1.0, kernel_name="cubic", eta=1.2)
    p = (2.0 / 3.0) * d.rho
    f = osph.forces_numba(pos, np.zeros_like(pos), mass, d.h, d.rho, d.omega, p, d.neighbours, "cubic")
    check("lattice: max |a| / (P / rho h)", 0.0, np.abs(f.acc).max() / (p[0] / (d.rho[0] * d.h[0])), atol=1e-12)
    check("lattice at rest: du/dt = 0", 0.0, np.abs(f.dudt).max(), atol=1e-14)


def test_conservation(rng):
    pos, vel, mass, d, p, u, cs, alpha, nb = _moving(rng, 2000)
    f = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", u=u, c_s=cs, alpha=alpha)
    mom = np.sum(mass[:, None] * f.acc, axis=0)
    check("sum m a = 0", 0.0, np.linalg.norm(mom) / np.sum(mass * np.linalg.norm(f.acc, axis=1)), atol=1e-13)
    torque = np.sum(mass[:, None] * np.cross(pos, f.acc), axis=0)
    ts = np.sum(mass * np.linalg.norm(pos, axis=1) * np.linalg.norm(f.acc, axis=1))
    check("sum m x cross a = 0", 0.0, np.linalg.norm(torque) / ts, atol=1e-13)
    power = np.sum(mass * (np.sum(vel * f.acc, axis=1) + f.dudt))
    ps = np.sum(mass * (np.abs(np.sum(vel * f.acc, axis=1)) + np.abs(f.dudt)))
    check("d/dt sum m (v^2/2 + u) = 0", 0.0, abs(power) / ps, atol=1e-13)


def test_compression_heats_and_av_only_approaching():
    pos = fcc_lattice(10)
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 1.2 * (1.0 / n) ** (1 / 3)), kernel_name="cubic", eta=1.2)
    p = (2.0 / 3.0) * d.rho
    vel = -(pos - 0.5)
    f = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, d.neighbours, "cubic")
    inner = np.linalg.norm(pos - 0.5, axis=1) < 0.25
    check("contraction: du/dt = 3 P / rho", 3 * p[inner] / d.rho[inner], f.dudt[inner], rtol=1e-2)
    cs = np.full(n, 0.1)
    al = np.ones(n)
    rec = osph.forces_numba(pos, -vel, mass, d.h, d.rho, d.omega, p, d.neighbours, "cubic", c_s=cs, alpha=al)
    check("receding flow: no viscous heating", 0.0, np.abs(rec.dudt_visc).max())
    app = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, d.neighbours, "cubic", c_s=cs, alpha=al)
    check("approaching flow: viscous heating > 0", True, bool(np.all(app.dudt_visc[inner] > 0)))


def test_capped_pair_fades():
    h_max, m = 200.0, np.full(2, 1.0 / 50015)
    acc = {}
    for sep in (300.0, 395.0, 399.0):
        pos = np.array([[0.0, 0.0, 0.0], [sep, 0.0, 0.0]])
        d = osph.compute_density(pos, m, np.full(2, h_max), kernel_name="cubic", eta=1.2, h_max=h_max)
        check(f"capped pair r={sep}: Omega = 1", [1.0, 1.0], d.omega)
        f = osph.forces_numba(pos, np.zeros_like(pos), m, d.h, d.rho, d.omega, (2 / 3) * d.rho * 5e-3,
                              d.neighbours, "cubic")  # fmt: skip
        acc[sep] = np.linalg.norm(f.acc[0])
    check("push at 395 < 0.05 push at 300", True, acc[395.0] < 0.05 * acc[300.0])
    check("push at 399 < 0.1 push at 395", True, acc[399.0] < 0.1 * acc[395.0])


# ----------------------------------------------------------------- gradients / viscosity
def test_gradients_exact_for_linear_fields(rng):
    """The corrected gradient is exact for linear velocity fields (sph.py:693-712)."""
    pos = fcc_lattice(10) + rng.uniform(-0.01, 0.01, (4000, 3))
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 1.2 * (1.0 / n) ** (1 / 3)), kernel_name="cubic", eta=1.2)
    nb = on.tighten(d.neighbours, pos, d.h)
    G = rng.normal(0, 1, (3, 3))
    vel = pos @ G.T
    acc = pos @ (G @ G).T
    g = osph.gradients_numba(pos, vel, acc, mass, d.h, nb, "cubic")
    inner = np.all(np.abs(pos - 0.5) < 0.3, axis=1)
    check("div v = tr G", np.full(inner.sum(), np.trace(G)), g.div_v[inner], rtol=1e-10)
    c = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    check("|curl v| = |antisym G|", np.full(inner.sum(), np.linalg.norm(c)), g.curl_v[inner], rtol=1e-10)
    check("d div v/dt = tr(G^2) - sum G_ij G_ji = 0", 0.0, np.abs(g.ddivv_dt[inner]).max(), atol=1e-9)
    gv = osph.velocity_gradients_numba(pos, vel, mass, d.h, nb, "cubic")
    check("velocity gradient = G", np.broadcast_to(G, (inner.sum(), 3, 3)), gv[inner], atol=1e-10)
    # the uncorrected difference operator is only approximate: checked on the unjittered lattice
    pos = fcc_lattice(10)
    d = osph.compute_density(pos, mass, np.full(n, 1.2 * (1.0 / n) ** (1 / 3)), kernel_name="cubic", eta=1.2)
    nb = on.tighten(d.neighbours, pos, d.h)
    ns = osph.ns_gradients_numba(pos, pos @ G.T, mass, d.h, d.rho, d.omega, nb, "cubic")
    inner = np.all(np.abs(pos - 0.5) < 0.3, axis=1)
    err = np.linalg.norm(ns[inner] - G, axis=(1, 2)) / np.linalg.norm(G)
    check("NS difference-operator gradient of v = G x: median error < 3%", 0.0, float(np.median(err)), atol=0.03)


def test_disc_and_ns_conserve(rng):
    pos, vel, mass, d, p, u, cs, alpha, nb = _moving(rng, 1200)
    a, du = osph.disc_forces_numba(pos, vel, mass, d.h, d.rho, d.omega, cs, 0.5, nb, "cubic")
    ms = np.sum(mass * np.linalg.norm(a, axis=1))
    check("disc: sum m a = 0", 0.0, np.linalg.norm(np.sum(mass[:, None] * a, axis=0)) / ms, atol=1e-13)
    check("disc: sum m x cross a = 0", 0.0, np.linalg.norm(np.sum(mass[:, None] * np.cross(pos, a), axis=0))
          / np.sum(mass * np.linalg.norm(pos, axis=1) * np.linalg.norm(a, axis=1)), atol=1e-13)  # fmt: skip
    pw = np.sum(mass * (np.sum(vel * a, axis=1) + du))
    check("disc: energy conserved", 0.0, abs(pw) / np.sum(mass * np.abs(du)), atol=1e-12)
    check("disc: heating >= 0", True, bool(np.all(du >= -1e-15 * np.abs(du).max())))
    g = osph.ns_gradients_numba(pos, vel, mass, d.h, d.rho, d.omega, nb, "cubic")
    nu = np.full(len(pos), 0.01)
    st = oph.viscous_stress(g, d.rho, nu)
    an = osph.ns_forces_numba(pos, mass, d.h, d.rho, d.omega, st, nb, "cubic")
    heat = oph.viscous_heating(g, nu)
    pw = np.sum(mass * (np.sum(vel * an, axis=1) + heat))
    check("NS: sum m (v.a + Lambda) = 0 (Phantom Eq. 132)", 0.0, abs(pw) / np.sum(mass * np.abs(heat)), atol=1e-12)


def test_inverse_and_determinant_match_lapack(rng):
    """The gradient passes invert R with the oracle's own 3x3 elimination (the reference calls
    np.linalg.inv / det, i.e. LAPACK): both agree to round-off on well-conditioned and on
    badly scaled matrices, and a transposed or sign-flipped result would fail."""
    from oracle import lib

    n = 2000
    a = rng.normal(0.0, 1.0, (n, 3, 3)) + 3.0 * np.eye(3)
    a[: n // 4] *= rng.uniform(1e-6, 1e6, (n // 4, 1, 1))  # scale does not matter
    inv = np.empty_like(a)
    det = np.empty(n)
    lib.inv3_det3(lib.d(np.ascontiguousarray(a)), n, lib.d(inv), lib.d(det))
    ref_inv = np.linalg.inv(a)
    ref_det = np.linalg.det(a)
    rel = np.abs(inv - ref_inv).max(axis=(1, 2)) / np.abs(ref_inv).max(axis=(1, 2))
    check("3x3 inverse vs LAPACK (max rel.)", 0.0, float(rel.max()), atol=1e-13)
    check("3x3 determinant vs LAPACK", ref_det, det, rtol=1e-13)
    eye = np.einsum("nij,njk->nik", a, inv)
    check("A A^-1 = I", np.broadcast_to(np.eye(3), eye.shape), eye, atol=1e-12)


def test_balsara_closed_forms():
    """Balsara (1995) limiter on gradients of closed-form flows (the BASELINE C3 switch, which the
    reference lacks): pure compression keeps alpha (f -> 1 up to the eps c_s/h term), pure
    rotation removes it (f = 0 exactly on the perfect lattice)."""
    from oracle import physics as ph
    from oracle import sph as osph
    from oracle.neighbours import tighten

    k = 12
    g = (np.arange(k) + 0.5) / k
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    mass = np.full(n, 1.0 / n)
    d = osph.compute_density(pos, mass, np.full(n, 1.5 / k), box=1.0, kernel_name="cubic", eta=1.2)
    nb = tighten(d.neighbours, pos, d.h)
    cs = np.full(n, 1e-3)
    # compression v = -x about the centre (periodic images far from the probes)
    inner = np.linalg.norm(pos - 0.5, axis=1) < 0.25
    v = -(pos - 0.5)
    gr = osph.gradients_numba(pos, v, None, mass, d.h, nb, "cubic")
    f = ph.balsara_factor(gr.div_v, gr.curl_v, cs, d.h)
    check("compression: div v = -3", -3.0 * np.ones(inner.sum()), gr.div_v[inner], rtol=1e-10)
    expect = 3.0 / (3.0 + ph.BALSARA_EPS * cs[inner] / d.h[inner])
    check("compression: f = 3 / (3 + eps c_s/h)", expect, f[inner], rtol=1e-9)
    # rigid rotation about z: v = omega x (x - c)
    w = np.array([0.0, 0.0, 2.0])
    v = np.cross(w, pos - 0.5)
    gr = osph.gradients_numba(pos, v, None, mass, d.h, nb, "cubic")
    f = ph.balsara_factor(gr.div_v, gr.curl_v, cs, d.h)
    check("rotation: |curl v| = 2 |omega|", np.full(inner.sum(), 4.0), gr.curl_v[inner], rtol=1e-10)
    check("rotation: f = 0", np.zeros(inner.sum()), f[inner], atol=1e-12)
