"""SPH passes, oracle side (test infrastructure only).

Host wrappers of the C restatements of sph.py's numba kernels, with the reference's
argument meaning. Flags: 0 ok, 1 at h_max, 2 search too small, 3 no convergence
(sph.py:36-39).
"""

from dataclasses import dataclass

import numpy as np

from . import lib
from .neighbours import Neighbours, find_neighbours_grid, minimum_image

KERNEL_IDS = {"cubic": 0, "wendland_c4": 1}  # kernels_nb.py:12
C_NORM = np.array([1.0 / np.pi, 495.0 / (256.0 * np.pi)])  # kernels.py:94-95
FLAG_OK, FLAG_AT_H_MAX, FLAG_SEARCH_TOO_SMALL, FLAG_NO_CONVERGENCE = 0, 1, 2, 3


@dataclass
class DensityResult:
    h: np.ndarray
    rho: np.ndarray
    omega: np.ndarray
    iterations: np.ndarray
    flags: np.ndarray
    neighbours: Neighbours
    builds: int = 1
    cache: object = None


@dataclass
class NeighbourCache:
    """sph.py:62-82: lists built at x_ref with search lengths h_search."""

    neighbours: Neighbours
    x_ref: np.ndarray
    h_search: np.ndarray
    builds: int = 1
    reuses: int = 0


FAST_SET_SIZES = (0, 1, 2, 4, 8, 16, 32, 64)  # sph.py:114


def fast_rows_complete(pos, fast, h_lim, box_arr, nb: Neighbours) -> bool:
    """sph.py:86-111: for every j in fast, every k with r_jk <= 2 max(h_lim_j, h_lim_k) at pos is
    in j's list (brute force over all k; per-component squares summed in order x, y, z)."""
    pos = np.asarray(pos, float)
    for j in fast:
        d = minimum_image(pos - pos[j], None if box_arr is None else np.asarray(box_arr, float))
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        rr = 2.0 * np.maximum(h_lim[j], h_lim)
        need = np.nonzero(d2 <= rr * rr)[0]
        if not np.all(np.isin(need, nb.of(int(j)))):
            return False
    return True


def reuse_limits(cache: NeighbourCache, pos, h_guess, rows, others, box_arr):
    """sph.py:117-147: h_eff with which the cached lists are provably complete, or None. Moves
    are ordered by size, ties by index (the reference leaves the order of ties to argpartition;
    a tie changes no h_eff, only which of the equal movers is checked directly)."""
    d = minimum_image(np.asarray(pos, float) - cache.x_ref, None if box_arr is None else np.asarray(box_arr, float))
    disp = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    n = len(disp)
    top = min(n, FAST_SET_SIZES[-1] + 1)
    order = np.lexsort((np.arange(n), -disp))[:top]
    for nf in FAST_SET_SIZES:
        if nf >= len(order):
            break
        h_eff = cache.h_search - 0.5 * (disp + disp[order[nf]])
        if not (np.all(h_eff[rows] > 0.5 * h_guess[rows]) and np.all(h_guess[others] <= h_eff[others])):
            continue
        if nf == 0:
            return h_eff
        h_lim = h_eff.copy()
        h_lim[others] = h_guess[others]
        if fast_rows_complete(pos, order[:nf], h_lim, box_arr, cache.neighbours):
            return h_eff
    return None


@dataclass
class ForceResult:
    acc: np.ndarray
    dudt: np.ndarray
    vsig_max: np.ndarray
    dudt_visc: np.ndarray
    dudt_cond: np.ndarray


@dataclass
class Gradients:
    div_v: np.ndarray
    curl_v: np.ndarray
    ddivv_dt: np.ndarray


def _rows(rows, n):
    return np.arange(n, dtype=np.int64) if rows is None else lib.i64(rows)


def _box(nb):
    return np.zeros(3) if nb.box is None else lib.f64(nb.box)


def kernel_f(q, kid):
    """kernels_nb.py:16-28"""
    return lib.kernel_f_c(float(q), int(kid))


def kernel_df(q, kid):
    """kernels_nb.py:31-43"""
    return lib.kernel_df_c(float(q), int(kid))


def density_numba(pos, mass, h_start, nb: Neighbours, h_search, kernel_name="cubic", eta=1.2,
                  tol=1e-4, h_max=np.inf, rows=None) -> DensityResult:  # fmt: skip
    """sph.py:300-326 -> _density_kernel (sph.py:251-297)."""
    rows = _rows(rows, len(pos))
    n = len(rows)
    h_out, rho, omega = np.empty(n), np.empty(n), np.empty(n)
    iters, flags = np.empty(n, np.int64), np.empty(n, np.int64)
    kid = KERNEL_IDS[kernel_name]
    pos, mass, h_start, h_search = (lib.f64(a) for a in (pos, mass, h_start, h_search))
    lib.density(lib.d(pos), lib.d(mass), lib.d(h_start), lib.d(h_search), lib.i(nb.offset),
                lib.i(nb.index), lib.d(_box(nb)), kid, C_NORM[kid], float(eta), float(tol),
                float(h_max), lib.i(rows), n, lib.d(h_out), lib.d(rho), lib.d(omega),
                lib.i(iters), lib.i(flags))  # fmt: skip
    return DensityResult(h_out, rho, omega, iters, flags, nb)


def compute_density(pos, mass, h_guess, box=None, kernel_name="cubic", eta=1.2, tol=1e-4,
                    h_max=np.inf, skin=1.4, rows=None, cache=None) -> DensityResult:  # fmt: skip
    """sph.py:329-387: lists reused from ``cache`` while reuse_limits proves them complete,
    else built at search 2 min(h_guess, h_max) skin (capped at a quarter periodic box); flagged
    rows are redone with doubled search, up to 12 rounds. The result carries the cache
    (``builds`` counts the builds since the first, as the reference's NeighbourCache does)."""
    h_guess = np.asarray(h_guess, float)
    n_all = len(pos)
    rows = _rows(rows, n_all)
    others = np.ones(n_all, bool)
    others[rows] = False
    box_arr = None if box is None else np.broadcast_to(np.asarray(box, float), (3,))
    box_cap = np.inf if box_arr is None else 0.2499 * float(np.min(box_arr))
    if cache is not None and len(cache.x_ref) == n_all and n_all:
        h_eff = reuse_limits(cache, pos, h_guess, rows, others, box_arr)
        if h_eff is not None:
            res = density_numba(pos, mass, h_guess, cache.neighbours, h_eff, kernel_name, eta, tol, h_max, rows)
            if not np.any(res.flags == FLAG_SEARCH_TOO_SMALL):
                cache.reuses += 1
                res.cache, res.builds = cache, cache.builds
                return res
            h_guess = h_guess.copy()
            h_guess[rows] = np.maximum(h_guess[rows], res.h)
    h_search = np.minimum(np.minimum(h_guess, h_max) * skin, box_cap)
    builds = 0 if cache is None else cache.builds
    for _ in range(12):
        nb = find_neighbours_grid(pos, h_search, box=box_arr)
        builds += 1
        res = density_numba(pos, mass, h_guess, nb, h_search, kernel_name, eta, tol, h_max, rows)
        grow = res.flags == FLAG_SEARCH_TOO_SMALL
        if not np.any(grow):
            res.builds = builds
            reuses = 0 if cache is None else cache.reuses
            res.cache = NeighbourCache(nb, np.array(pos, float), h_search.copy(), builds, reuses)
            return res
        h_search = h_search.copy()
        h_search[rows[grow]] = np.minimum(h_search[rows[grow]] * 2.0, box_cap)
        h_guess = h_guess.copy()
        h_guess[rows] = res.h
    raise RuntimeError("smoothing lengths kept outgrowing the neighbour search")


def forces_numba(pos, vel, mass, h, rho, omega, pressure, nb: Neighbours, kernel_name="cubic",
                 u=None, c_s=None, alpha=None, beta=2.0, alpha_u=1.0, cond_signal=0,
                 reconstruction=False, closeness=True, eta_crit=1.0 / 1.2, eta_fold=0.2,
                 recon_gate=2, rows=None, grad_v=None) -> ForceResult:  # fmt: skip
    """sph.py:655-690 -> _forces_kernel (sph.py:525-652). Viscosity iff alpha and c_s are
    given; conductivity iff u is given as well (sph.py:668-670)."""
    n = len(pos)
    rows = _rows(rows, n)
    r = len(rows)
    acc, dudt, vsig = np.empty((r, 3)), np.empty(r), np.empty(r)
    vheat, cheat = np.empty(r), np.empty(r)
    kid = KERNEL_IDS[kernel_name]
    use_visc = alpha is not None and c_s is not None
    use_cond = use_visc and u is not None
    use_recon = use_visc and reconstruction
    if use_recon and (grad_v is None or np.shape(grad_v) != (n, 3, 3)):
        raise ValueError("velocity reconstruction needs grad_v of shape (N, 3, 3)")
    zeros = np.zeros(n)
    gv = lib.f64(grad_v) if use_recon else np.zeros((1, 3, 3))

    def arr(x):
        return zeros if x is None else lib.f64(x)

    lib.forces(lib.d(lib.f64(pos)), lib.d(lib.f64(vel)), lib.d(lib.f64(mass)), lib.d(lib.f64(h)),
               lib.d(lib.f64(rho)), lib.d(lib.f64(omega)), lib.d(lib.f64(pressure)), lib.d(arr(u)),
               lib.d(arr(c_s)), lib.d(arr(alpha)), lib.i(nb.offset), lib.i(nb.index),
               lib.d(_box(nb)), kid, C_NORM[kid], int(use_visc), int(use_cond), float(beta),
               float(alpha_u), int(cond_signal), lib.i(rows), r, int(use_recon), int(closeness),
               float(eta_crit), float(eta_fold), int(recon_gate), lib.d(gv), lib.d(acc),
               lib.d(dudt), lib.d(vsig), lib.d(vheat), lib.d(cheat))  # fmt: skip
    return ForceResult(acc, dudt, vsig, vheat, cheat)


def gradients_numba(pos, vel, acc, mass, h, nb: Neighbours, kernel_name="cubic",
                    rows=None) -> Gradients:  # fmt: skip
    """sph.py:887-902 -> _gradients_kernel (sph.py:837-884). ``acc`` None means zeros.
    The 3x3 inverse/determinant are this oracle's own (Gaussian elimination with partial
    pivoting); the reference takes them from LAPACK, so bitwise parity with the reference
    is unpinned for rows that are not near-singular -- closed-form tests pin the result."""
    rows = _rows(rows, len(pos))
    r = len(rows)
    div, curl, ddiv = np.empty(r), np.empty(r), np.empty(r)
    kid = KERNEL_IDS[kernel_name]
    lib.gradients(lib.d(lib.f64(pos)), lib.d(lib.f64(vel)),
                  None if acc is None else lib.d(lib.f64(acc)), lib.d(lib.f64(mass)),
                  lib.d(lib.f64(h)), lib.i(nb.offset), lib.i(nb.index), lib.d(_box(nb)), kid,
                  C_NORM[kid], lib.i(rows), r, lib.d(div), lib.d(curl), lib.d(ddiv))  # fmt: skip
    return Gradients(div, curl, ddiv)


def velocity_gradients_numba(pos, vel, mass, h, nb: Neighbours, kernel_name="cubic", rows=None,
                             out=None) -> np.ndarray:  # fmt: skip
    """sph.py:818-834 -> _velocity_gradient_kernel (sph.py:785-815); (N, 3, 3)."""
    n = len(pos)
    rows = _rows(rows, n)
    out = np.zeros((n, 3, 3)) if out is None else out
    kid = KERNEL_IDS[kernel_name]
    lib.velocity_gradients(lib.d(lib.f64(pos)), lib.d(lib.f64(vel)), lib.d(lib.f64(mass)),
                           lib.d(lib.f64(h)), lib.i(nb.offset), lib.i(nb.index), lib.d(_box(nb)),
                           kid, C_NORM[kid], lib.i(rows), len(rows), lib.d(out))  # fmt: skip
    return out


def ns_gradients_numba(pos, vel, mass, h, rho, omega, nb, kernel_name="cubic", rows=None,
                       out=None):  # fmt: skip
    """sph.py:984-999 -> _ns_gradient_kernel (sph.py:957-981); (N, 3, 3)."""
    n = len(pos)
    rows = _rows(rows, n)
    out = np.zeros((n, 3, 3)) if out is None else out
    kid = KERNEL_IDS[kernel_name]
    lib.ns_gradients(lib.d(lib.f64(pos)), lib.d(lib.f64(vel)), lib.d(lib.f64(mass)),
                     lib.d(lib.f64(h)), lib.d(lib.f64(rho)), lib.d(lib.f64(omega)),
                     lib.i(nb.offset), lib.i(nb.index), lib.d(_box(nb)), kid, C_NORM[kid],
                     lib.i(rows), len(rows), lib.d(out))  # fmt: skip
    return out


def ns_forces_numba(pos, mass, h, rho, omega, stress, nb, kernel_name="cubic", rows=None):
    """sph.py:1030-1045 -> _ns_force_kernel (sph.py:1002-1027); (R, 3)."""
    rows = _rows(rows, len(pos))
    acc = np.empty((len(rows), 3))
    kid = KERNEL_IDS[kernel_name]
    lib.ns_forces(lib.d(lib.f64(pos)), lib.d(lib.f64(mass)), lib.d(lib.f64(h)),
                  lib.d(lib.f64(rho)), lib.d(lib.f64(omega)), lib.d(lib.f64(stress)),
                  lib.i(nb.offset), lib.i(nb.index), lib.d(_box(nb)), kid, C_NORM[kid],
                  lib.i(rows), len(rows), lib.d(acc))  # fmt: skip
    return acc


def disc_forces_numba(pos, vel, mass, h, rho, omega, c_s, alpha_av, nb, kernel_name="cubic",
                      rows=None):  # fmt: skip
    """sph.py:1120-1137 -> _disc_kernel (sph.py:1084-1117); (acc (R,3), dudt (R,))."""
    rows = _rows(rows, len(pos))
    acc, dudt = np.empty((len(rows), 3)), np.empty(len(rows))
    kid = KERNEL_IDS[kernel_name]
    lib.disc_forces(lib.d(lib.f64(pos)), lib.d(lib.f64(vel)), lib.d(lib.f64(mass)),
                    lib.d(lib.f64(h)), lib.d(lib.f64(rho)), lib.d(lib.f64(omega)),
                    lib.d(lib.f64(c_s)), float(alpha_av), lib.i(nb.offset), lib.i(nb.index),
                    lib.d(_box(nb)), kid, C_NORM[kid], lib.i(rows), len(rows), lib.d(acc),
                    lib.d(dudt))  # fmt: skip
    return acc, dudt
