"""GPU parity of every HIP pass against the FP64 CPU oracle on identical seeded inputs.

FP64 passes are compared bit for bit (same operations, same order, IEEE division and square
root, no contraction). The velocity-reconstruction path calls exp(), whose last bit may differ
between the device library and glibc: it is compared at 1e-12 relative.
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _uneven(rng, n, box):
    """test_neighbours.py:15-19 of the reference: h varies 30x so gather-only lists miss pairs."""
    pos = rng.uniform(0, box, size=(n, 3))
    h = np.exp(rng.uniform(np.log(0.01), np.log(0.3), n)) * box / 8
    return pos, h


def blob(rng, n=2000):
    """test_density.py:21-27 of the reference: Gaussian clump in a sparse background."""
    core = rng.normal(0.0, 0.15, size=(n // 2, 3))
    bg = rng.uniform(-1.0, 1.0, size=(n - n // 2, 3))
    return np.concatenate([core, bg]), np.full(n, 1.0 / n)


def _geometry(rng, case):
    if case == "open_uneven":
        pos, h = _uneven(rng, 3000, 2.0)
        return pos, h, None
    if case == "periodic_cube":
        pos, h = _uneven(rng, 3000, 2.0)
        return pos, h, 2.0
    if case == "periodic_long_box":
        box = np.array([2.0, 0.125, 0.125])
        return rng.uniform(0, 1, (4000, 3)) * box, rng.uniform(0.005, 0.03, 4000), box
    if case == "isolated_large_h":
        pos = np.concatenate([rng.uniform(0, 1, (2000, 3)), [[5.0, 5.0, 5.0], [-3.0, 0.5, 0.5]]])
        return pos, np.concatenate([np.full(2000, 0.04), [4.0, 2.5]]), None
    pos = np.concatenate([rng.normal(0, 0.02, (2000, 3)), rng.uniform(-1, 1, (500, 3))])
    return pos, np.concatenate([np.full(2000, 0.005), np.full(500, 0.5)]), None


@pytest.mark.parametrize(
    "case", ["open_uneven", "periodic_cube", "periodic_long_box", "isolated_large_h", "clump_and_halo"]
)
def test_lists_bit_identical(rng, case):
    from nebula.sim.neighbours import find_neighbours_numba

    from oracle.neighbours import find_neighbours_grid

    pos, h, box = _geometry(rng, case)
    print(f"lists case={case} n={len(pos)}", flush=True)
    ref = find_neighbours_grid(pos, h, box=box)
    got = find_neighbours_numba(pos, h, box=box)
    bit_equal(f"{case}: offsets", ref.offset, got.offset)
    bit_equal(f"{case}: indices", ref.index, got.index)


def _kdtree_lists(pos, h, box=None):
    """The reference's cKDTree twin (neighbours.py:53-71) restated with scipy: query_ball_point
    at r = 2 h (boundary included), made symmetric, rows sorted."""
    import scipy.sparse as sp
    from scipy.spatial import cKDTree

    n = len(pos)
    box_arr = None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()
    w = pos if box_arr is None else np.mod(pos, box_arr)
    lists = cKDTree(w, boxsize=box_arr).query_ball_point(w, r=2.0 * h, return_sorted=False)
    lengths = np.array([len(x) for x in lists], np.int64)
    cols = np.fromiter((j for x in lists for j in x), np.int64, int(lengths.sum()))
    g = sp.csr_matrix((np.ones(len(cols), np.int8), (np.repeat(np.arange(n), lengths), cols)), shape=(n, n))
    sym = (g + g.T).tocsr()
    sym.sort_indices()
    return sym.indptr.astype(np.int64), sym.indices.astype(np.int64)


@pytest.mark.parametrize("case", ["lattice_boundary", "open_uneven", "periodic_cube"])
def test_find_neighbours_is_the_inclusive_kdtree_twin(rng, case):
    """find_neighbours keeps pairs at exactly 2 max(h) (cKDTree's query_ball_point includes its
    boundary); find_neighbours_numba's grid search does not. On a unit lattice with h = 0.5 every
    nearest-neighbour pair sits on that boundary."""
    from nebula.sim.neighbours import find_neighbours, find_neighbours_numba

    if case == "lattice_boundary":
        g = np.arange(6, dtype=np.float64)
        pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        h, box = np.full(len(pos), 0.5), None
    else:
        pos, h, box = _geometry(rng, case)
    print(f"find_neighbours case={case} n={len(pos)}", flush=True)
    off, idx = _kdtree_lists(pos, h, box)
    got = find_neighbours(pos, h, box=box)
    bit_equal(f"{case}: offsets = the cKDTree twin's", off, got.offset)
    bit_equal(f"{case}: indices = the cKDTree twin's", idx, got.index)
    check(f"{case}: host index is int64 as in the reference", np.dtype(np.int64), got.index.dtype)
    if case == "lattice_boundary":
        strict = find_neighbours_numba(pos, h)
        check("boundary pairs: only self in the strict lists", len(pos), int(strict.offset[-1]))
        check("boundary pairs: 6 lattice neighbours + self inside", 7, int(np.max(np.diff(got.offset))))


def test_single_particle_and_empty():
    from nebula.sim.neighbours import find_neighbours_numba

    nb = find_neighbours_numba(np.array([[0.3, 0.2, 0.1]]), np.array([0.1]))
    check("single particle lists itself only", [0], nb.of(0).tolist())
    nb0 = find_neighbours_numba(np.zeros((0, 3)), np.zeros(0))
    check("no particles: empty lists", 0, int(nb0.offset[-1]))


def test_periodic_box_too_small_raises():
    from nebula.sim.neighbours import find_neighbours_numba

    with pytest.raises(ValueError, match="half the periodic box"):
        find_neighbours_numba(np.zeros((2, 3)) + 0.5, np.array([0.3, 0.3]), box=1.0)


def test_tighten_bit_identical(rng):
    from nebula.sim.neighbours import find_neighbours_numba, tighten

    from oracle import neighbours as on

    pos, h = _uneven(rng, 3000, 2.0)
    loose_ref = on.find_neighbours_grid(pos, 1.25 * h, box=2.0)
    loose = find_neighbours_numba(pos, 1.25 * h, box=2.0)
    ref = on.tighten(loose_ref, pos, h)
    got = tighten(loose, pos, h)
    bit_equal("tighten offsets", ref.offset, got.offset)
    bit_equal("tighten indices", ref.index, got.index)
    rows = np.arange(0, 3000, 7)
    ref_r = on.tighten(loose_ref, pos, h, rows=rows)
    got_r = tighten(loose, pos, h, rows=rows)
    bit_equal("tighten(rows) offsets", ref_r.offset, got_r.offset)
    bit_equal("tighten(rows) indices", ref_r.index, got_r.index)


@pytest.mark.parametrize("kernel", ["cubic", "wendland_c4"])
def test_density_bit_identical(rng, kernel):
    from nebula.sim.sph import compute_density, density_numba

    from oracle import neighbours as on
    from oracle import sph as osph

    pos, mass = blob(rng)
    eta = 1.2 if kernel == "cubic" else 1.6
    h0 = np.full(len(pos), 0.05)
    ref = osph.compute_density(pos, mass, h0, kernel_name=kernel, eta=eta, h_max=0.5)
    got = compute_density(pos, mass, h0, kernel_name=kernel, eta=eta, h_max=0.5)
    for k in ("h", "rho", "omega", "iterations", "flags"):
        bit_equal(f"{kernel} compute_density {k}", getattr(ref, k), getattr(got, k))
    # given lists, a row subset, non-uniform masses
    mass2 = mass * rng.uniform(0.5, 1.5, len(mass))
    hs = ref.h * 1.25
    nb_ref = on.find_neighbours_grid(pos, hs)
    rows = np.arange(3, len(pos), 5)
    r = osph.density_numba(pos, mass2, ref.h, nb_ref, hs, kernel, eta, 1e-12, 0.5, rows)
    from nebula.sim.neighbours import Neighbours

    g = density_numba(pos, mass2, ref.h, Neighbours(nb_ref.offset, nb_ref.index, None), hs, kernel, eta, 1e-12,
                      0.5, rows)  # fmt: skip
    for k in ("h", "rho", "omega", "iterations", "flags"):
        bit_equal(f"{kernel} density_numba(rows, tol 1e-12) {k}", getattr(r, k), getattr(g, k))


def test_density_isolated_particle_flag():
    from nebula.sim.sph import FLAG_AT_H_MAX, compute_density

    a = 1.0 / 8
    g = (np.arange(8) + 0.25) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([pos[pos[:, 0] < 0.5], [[5.0, 5.0, 5.0]]])
    n = len(pos)
    res = compute_density(pos, np.full(n, 1 / 512), np.full(n, 1.2 * a), kernel_name="cubic", eta=1.2, h_max=0.4)
    check("isolated particle flagged at h_max", FLAG_AT_H_MAX, int(res.flags[-1]))
    check("isolated particle h = h_max", 0.4, res.h[-1], rtol=1e-12)
    check("no NaN", False, bool(np.any(np.isnan(res.rho) | np.isnan(res.omega))))


def _moving_state(rng, kernel="cubic"):
    from oracle import sph as osph
    from oracle.neighbours import tighten

    pos, mass = blob(rng)
    eta = 1.2 if kernel == "cubic" else 1.6
    d = osph.compute_density(pos, mass, np.full(len(pos), 0.05), kernel_name=kernel, eta=eta, h_max=0.5)
    u = rng.uniform(0.5, 1.5, len(pos))
    vel = rng.normal(0.0, 0.3, size=pos.shape)
    p = (2.0 / 3.0) * d.rho * u
    cs = np.sqrt(10.0 / 9.0 * u)
    alpha = rng.uniform(0, 1, len(pos))
    nb = tighten(d.neighbours, pos, d.h)
    return pos, vel, mass, d, p, u, cs, alpha, nb


@pytest.mark.parametrize("mode", ["pressure", "visc", "visc_cond_p", "visc_cond_v"])
def test_forces_bit_identical(rng, mode):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import ShockCapturing, forces_numba

    from oracle import sph as osph

    pos, vel, mass, d, p, u, cs, alpha, nb = _moving_state(rng)
    use_visc = mode != "pressure"
    use_u = mode.startswith("visc_cond")
    sig = 1 if mode.endswith("_v") else 0
    kw = dict(u=u if use_u else None, c_s=cs if use_visc else None, alpha=alpha if use_visc else None)
    ref = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", beta=2.0, alpha_u=0.05,
                            cond_signal=sig, **kw)  # fmt: skip
    shock = ShockCapturing(beta=2.0, alpha_u=0.05, conductivity_signal=sig)
    got = forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, Neighbours(nb.offset, nb.index, None), "cubic",
                       shock=shock, **kw)  # fmt: skip
    bit_equal(f"{mode}: acc", ref.acc, got.acc)
    bit_equal(f"{mode}: dudt", ref.dudt, got.dudt)
    bit_equal(f"{mode}: vsig_max", ref.vsig_max, got.vsig_max)
    bit_equal(f"{mode}: dudt_visc", ref.dudt_visc, got.dudt_visc)
    bit_equal(f"{mode}: dudt_cond", ref.dudt_cond, got.dudt_cond)


@pytest.mark.parametrize("gate", [0, 1, 2])
def test_forces_reconstruction(rng, gate):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import ShockCapturing, forces_numba, velocity_gradients_numba

    from oracle import sph as osph

    pos, vel, mass, d, p, u, cs, alpha, nb = _moving_state(rng)
    gv_ref = osph.velocity_gradients_numba(pos, vel, mass, d.h, nb, "cubic")
    nbg = Neighbours(nb.offset, nb.index, None)
    gv = velocity_gradients_numba(pos, vel, mass, d.h, nbg, "cubic")
    bit_equal("velocity gradients", gv_ref, gv)
    ref = osph.forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nb, "cubic", u=u, c_s=cs, alpha=alpha,
                            alpha_u=0.05, cond_signal=1, reconstruction=True, recon_gate=gate, grad_v=gv_ref)
    shock = ShockCapturing(alpha_u=0.05, conductivity_signal=1, reconstruction=True, eta_crit=1 / 1.2,
                           recon_gate=gate)  # fmt: skip
    got = forces_numba(pos, vel, mass, d.h, d.rho, d.omega, p, nbg, "cubic", u=u, c_s=cs, alpha=alpha,
                       shock=shock, grad_v=gv)  # fmt: skip
    check(f"recon gate {gate}: acc (1e-12 per vector)", ref.acc, got.acc, rtol=1e-12, per_vector=True)
    s = np.abs(ref.dudt).max()
    check(f"recon gate {gate}: dudt (1e-12 of max)", ref.dudt / s, got.dudt / s, atol=1e-12)


def test_gradients_ns_disc_bit_identical(rng):
    from nebula.sim.neighbours import Neighbours
    from nebula.sim.sph import disc_forces_numba, gradients_numba, ns_forces_numba, ns_gradients_numba

    from oracle import physics as oph
    from oracle import sph as osph

    pos, vel, mass, d, p, u, cs, alpha, nb = _moving_state(rng)
    acc = rng.normal(0, 1, pos.shape)
    nbg = Neighbours(nb.offset, nb.index, None)
    for a in (acc, None):
        ref = osph.gradients_numba(pos, vel, a, mass, d.h, nb, "cubic")
        got = gradients_numba(pos, vel, a, mass, d.h, nbg, "cubic")
        for k in ("div_v", "curl_v", "ddivv_dt"):
            bit_equal(f"gradients {k} (acc {'given' if a is not None else 'zero'})", getattr(ref, k),
                      getattr(got, k))  # fmt: skip
    g_ref = osph.ns_gradients_numba(pos, vel, mass, d.h, d.rho, d.omega, nb, "cubic")
    g = ns_gradients_numba(pos, vel, mass, d.h, d.rho, d.omega, nbg, "cubic")
    bit_equal("NS gradients", g_ref, g)
    nu = oph.ss_viscosity(0.1, cs, np.linalg.norm(pos, axis=1) + 1.0)
    stress = oph.viscous_stress(g_ref, d.rho, nu)
    bit_equal("NS forces", osph.ns_forces_numba(pos, mass, d.h, d.rho, d.omega, stress, nb, "cubic"),
              ns_forces_numba(pos, mass, d.h, d.rho, d.omega, stress, nbg, "cubic"))
    ra, ru = osph.disc_forces_numba(pos, vel, mass, d.h, d.rho, d.omega, cs, 0.3, nb, "cubic")
    ga, gu = disc_forces_numba(pos, vel, mass, d.h, d.rho, d.omega, cs, 0.3, nbg, "cubic")
    bit_equal("disc acc", ra, ga)
    bit_equal("disc dudt", ru, gu)


def test_glf_and_v_from_p_bit_identical(rng):
    from nebula.sim.fastphys import gn_v_from_p_nb
    from nebula.sim.integrator import GLFConvergenceError, glf_step_many

    from oracle import integrator as oi
    from oracle import physics as oph

    n = 5000
    r = rng.uniform(4.0, 3000.0, n)
    u = rng.normal(size=(n, 3))
    x = u / np.linalg.norm(u, axis=1)[:, None] * r[:, None]
    v = rng.normal(0, 1, (n, 3)) * np.sqrt(1.0 / r)[:, None] * 0.5
    p = oph.gn_p_from_v(x, v)
    dt = 0.005 * r**1.5
    for dtv in (dt, float(np.min(dt))):
        x1r, p1r, itr = oi.glf_step_many(x, p, dtv)
        x1, p1, it = glf_step_many(x, p, dtv)
        bit_equal("GLF x1", x1r, x1)
        bit_equal("GLF p1", p1r, p1)
        bit_equal("GLF iterations", itr, it)
    bit_equal("v from p", oi.gn_v_from_p(x, p), gn_v_from_p_nb(x, p))
    xx = np.array([[6.0, 0.0, 0.0]])
    pp = oph.gn_p_from_v(xx, np.array([[0.0, 0.4, 0.0]]))
    with pytest.raises(GLFConvergenceError, match="did not converge"):
        glf_step_many(xx, pp, 50.0)
    with pytest.raises(GLFConvergenceError, match="inside r"):
        glf_step_many(np.array([[2.4, 0.0, 0.0]]), pp, 0.1)
