"""SPH passes on the GPU behind the reference's call surface (sph.py).

Density and smoothing length (grad-h SPH, Newton-Raphson with a bisection bracket), the
force pass (pressure, grad-h, Monaghan viscosity, conductivity, P dV heating), velocity
gradients, Navier-Stokes and disc viscosity: each function keeps the reference's name,
arguments, flags and result types (sph.py:36-1137) and runs one HIP kernel
(csrc/pairs_fp64.hip). NumPy inputs give NumPy results; torch GPU tensors give GPU tensors.
"""

from __future__ import annotations

from dataclasses import dataclass

import ctypes

import numpy as np
import torch

from .. import _dev, _hip
from ..physics.kernels import C_NORM, DEFAULT_ETA, DEFAULT_KERNEL, KERNEL_IDS
from .neighbours import Neighbours, build_lists, build_lists_blocked

FLAG_OK = 0
FLAG_AT_H_MAX = 1
FLAG_SEARCH_TOO_SMALL = 2
FLAG_NO_CONVERGENCE = 3
MAX_ITER = 50

COND_PRESSURE = 0
COND_VELOCITY = 1
GATE_BOTH, GATE_PUBLISHED, GATE_MINMOD = 0, 1, 2

PRECISIONS = {"fp64": 0, "mixed": 1}
# FP64 density passes of particles stored in grid order take the lane-group gather
# (nb_hydro_density64; same bits as nb_hydro_density, which the parity tests run beside it)
FP64_LANE_LISTS = True
NB_NOT_GRID_ORDER = 2


def _out(t, as_dev: bool):
    return t if as_dev else _dev.host(t)


@dataclass
class DensityResult:
    h: object
    rho: object
    omega: object
    iterations: object
    flags: object
    neighbours: Neighbours
    cache: object = None
    builds: int = 1


@dataclass
class NeighbourCache:
    """Neighbour lists kept between density passes ("Verlet lists", sph.py:62-82), on the
    device: built at positions ``x_ref`` with search lengths ``h_search``, so they held every pair
    with r < 2 max(hs_i, hs_j) then. ``compute_density`` reuses them while they provably still
    hold every interacting pair (``_reuse_limits``), and rebuilds them otherwise."""

    neighbours: Neighbours
    x_ref: object
    h_search: object
    builds: int = 1
    reuses: int = 0


FAST_SET_SIZES = (0, 1, 2, 4, 8, 16, 32, 64)  # sph.py:114


def reuse_limits_device(cache: NeighbourCache, pos_d, h_guess_d, role, box_arr):
    """Per-particle search limits h_eff (device) with which the cached lists are provably
    complete, or None when they must be rebuilt (sph.py:117-147). ``role`` (uint8 device, N):
    bit 0 marks the rows being solved, bit 1 the others (whose h stays h_guess).

    The moves d since the build are measured on the device; fast sets F of the largest movers
    are tried in growing sizes: h_eff = hs - (d + d_rest) / 2 with d_rest the largest move outside
    F; rows need h_eff > h_guess / 2, the others h_guess <= h_eff; the rows of F are then checked
    against every particle at the current positions (h_lim = h_eff for rows, h_guess for the
    others). Moves are ordered by size, ties by index: a tie cannot change h_eff, which takes the
    value of the move, only which of the equal movers is checked directly."""
    n = int(pos_d.shape[0])
    x_ref = _dev.f64(cache.x_ref)
    hs = _dev.f64(cache.h_search)
    disp = _dev.empty(n)
    _hip.call("nb_reuse_displacements", _hip.ptr(pos_d), _hip.ptr(x_ref), _hip.box_arg(box_arr), n, _hip.ptr(disp))
    top = min(n, FAST_SET_SIZES[-1] + 1)
    vals, order = torch.topk(disp, top, largest=True, sorted=True)
    vals_h, order_h = vals.cpu().numpy(), order.cpu().numpy()
    perm = np.lexsort((order_h, -vals_h))
    order_h, d_sorted = order_h[perm], vals_h[perm]
    h_eff, h_lim = _dev.empty(n), _dev.empty(n)
    bad = torch.zeros(2, dtype=torch.int32, device=pos_d.device)
    missing = torch.zeros(1, dtype=torch.int32, device=pos_d.device)
    nb = cache.neighbours
    for nf in FAST_SET_SIZES:
        if nf >= top:
            break
        bad.zero_()
        _hip.call("nb_reuse_limits", _hip.ptr(hs), _hip.ptr(disp), float(d_sorted[nf]), _hip.ptr(h_guess_d),
                  _hip.ptr(role), n, _hip.ptr(h_eff), _hip.ptr(h_lim), _hip.ptr(bad))  # fmt: skip
        if int(bad.sum().item()):
            continue
        if nf == 0:
            return h_eff
        fast = _dev.i64(order_h[:nf])
        missing.zero_()
        _hip.call("nb_reuse_fast_rows", _hip.ptr(pos_d), _hip.ptr(fast), nf, _hip.ptr(h_lim), _hip.box_arg(box_arr),
                  _hip.ptr(nb.d_offset), _hip.ptr(nb.d_index), n, _hip.ptr(missing))  # fmt: skip
        if int(missing.item()) == 0:
            return h_eff
    return None


def _roles(rows_d, others, n, device):
    role = torch.zeros(n, dtype=torch.uint8, device=device)
    if rows_d is None:
        role.fill_(1)
    else:
        role[rows_d] = 1
    if others is None:
        role[role == 0] = 2
    else:
        o = torch.as_tensor(np.asarray(others, bool) if not isinstance(others, torch.Tensor) else others,
                            device=device).to(torch.bool)  # fmt: skip
        role[o] |= 2
    return role


def _reuse_limits(cache: NeighbourCache, pos, h_guess, rows, others, box_arr):
    """sph.py:117-147 with the reference's arguments: ``rows`` the particles being solved,
    ``others`` a boolean mask of those whose h is fixed (normally the complement). Returns h_eff
    (host array for host positions, device tensor for device ones) or None."""
    as_dev = _dev.is_dev(pos)
    pos_d = _dev.f64(pos)
    n = int(pos_d.shape[0])
    rows_d, _ = _dev.rows_dev(rows, n)
    role = _roles(rows_d, others, n, pos_d.device)
    box = None if box_arr is None else np.broadcast_to(np.asarray(box_arr, np.float64), (3,)).copy()
    h_eff = reuse_limits_device(cache, pos_d, _dev.f64(h_guess), role, box)
    return None if h_eff is None else _out(h_eff, as_dev)


def _box3(box):
    return None if box is None else np.broadcast_to(np.asarray(box, np.float64), (3,)).copy()


def density_device(pos_d, mass_d, h_start_d, nb: Neighbours, h_search_d, kernel_name, eta, tol, h_max,
                   rows_d=None, nrows=None, m_uniform=None):  # fmt: skip
    """Device form of density_numba: returns (h, rho, omega, iters, flags) tensors."""
    n = int(pos_d.shape[0])
    nrows = n if nrows is None else nrows
    kid = KERNEL_IDS[kernel_name]
    if m_uniform is None:
        m_uniform = _dev.uniform_mass(mass_d)
    h_out, rho, omega = (_dev.empty(nrows) for _ in range(3))
    iters = _dev.empty(nrows, torch.int32)
    flags = _dev.empty(nrows, torch.int32)
    _hip.call("nb_density", _hip.ptr(pos_d), _hip.ptr(mass_d), float(m_uniform), _hip.ptr(h_start_d),
              _hip.ptr(h_search_d), ctypes.byref(nb.desc()), _hip.box_arg(nb.box), _hip.ptr(rows_d), nrows, kid, float(C_NORM[kid]), float(eta),
              float(tol), float(h_max), _hip.ptr(h_out), _hip.ptr(rho), _hip.ptr(omega),
              _hip.ptr(iters), _hip.ptr(flags))  # fmt: skip
    return h_out, rho, omega, iters, flags


def density_numba(pos, mass, h_start, nb: Neighbours, h_search, kernel_name=DEFAULT_KERNEL, eta=DEFAULT_ETA,
                  tol=1e-4, h_max=np.inf, rows=None) -> DensityResult:  # fmt: skip
    """The density pass on given neighbour lists (sph.py:300-326). ``rows``: solve only these
    particles (results in that order); default all."""
    as_dev = _dev.is_dev(pos)
    pos_d = _dev.f64(pos)
    rows_d, nrows = _dev.rows_dev(rows, int(pos_d.shape[0]))
    h, rho, omega, it, fl = density_device(pos_d, _dev.f64(mass), _dev.f64(h_start), nb, _dev.f64(h_search),
                                           kernel_name, eta, tol, h_max, rows_d, nrows)  # fmt: skip
    return DensityResult(*(_out(t, as_dev) for t in (h, rho, omega, it, fl)), nb)


def compute_density(pos, mass, h_guess, box=None, kernel_name=DEFAULT_KERNEL, eta=DEFAULT_ETA, tol=1e-4,
                    h_max=np.inf, skin=1.4, cache=None, rows=None) -> DensityResult:  # fmt: skip
    """Density pass with neighbour lists (sph.py:329-387): reused from ``cache`` while provably
    still complete (``_reuse_limits``), else built at search radius 2 min(h_guess, h_max) x skin;
    particles whose h outgrows the search are redone with rebuilt lists, doubling their search
    (up to 12 rounds). The result carries the cache for the next call. ``rows``: solve only these
    particles (results in that order); the lists stay complete for the others too."""
    as_dev = _dev.is_dev(pos)
    pos_d = _dev.f64(pos)
    mass_d = _dev.f64(mass)
    h_guess_d = _dev.f64(h_guess, copy=True)
    box_arr = _box3(box)
    n = int(pos_d.shape[0])
    res = None
    if cache is not None and len(cache.x_ref) == n and n:
        rows_d, nrows = _dev.rows_dev(rows, n)
        h_eff = reuse_limits_device(cache, pos_d, h_guess_d, _roles(rows_d, None, n, pos_d.device), box_arr)
        if h_eff is not None:
            h, rho, omega, it, fl = density_device(pos_d, mass_d, h_guess_d, cache.neighbours, h_eff, kernel_name,
                                                   eta, tol, h_max, rows_d, nrows)  # fmt: skip
            if not bool((fl == FLAG_SEARCH_TOO_SMALL).any().item()):
                cache.reuses += 1
                res = DensityResult(h, rho, omega, it, fl, cache.neighbours, cache, cache.builds)
            else:
                # rows whose h outgrew the reuse limit: the rebuild starts from at least their h
                if rows_d is None:
                    h_guess_d = torch.maximum(h_guess_d, h)
                else:
                    h_guess_d[rows_d] = torch.maximum(h_guess_d[rows_d], h)
    if res is None:
        res = compute_density_device(pos_d, mass_d, h_guess_d, box_arr, kernel_name, eta, tol, h_max, skin, rows,
                                     keep_search=True)  # fmt: skip
        prior = (0, 0) if cache is None else (cache.builds, cache.reuses)
        res.cache = NeighbourCache(res.neighbours, pos_d.clone(), res.cache, prior[0] + res.builds, prior[1])
        res.builds = res.cache.builds
    if not as_dev:
        for k in ("h", "rho", "omega", "iterations", "flags"):
            setattr(res, k, _dev.host(getattr(res, k)))
    return res


def compute_density_device(pos_d, mass_d, h_guess_d, box_arr, kernel_name, eta, tol, h_max, skin, rows=None,
                           m_uniform=None, capacity_hint=0, blocked=False, keep_search=False) -> DensityResult:  # fmt: skip
    """compute_density on device tensors (h_guess_d is overwritten by the rebuild rounds).
    ``blocked``: build the lists in 64-row blocks (the step's own layout) instead of CSR.
    ``keep_search``: the result's ``cache`` field holds the final search lengths (a
    NeighbourCache's h_search)."""
    n = int(pos_d.shape[0])
    rows_d, nrows = _dev.rows_dev(rows, n)
    box_cap = np.inf if box_arr is None else 0.2499 * float(np.min(box_arr))
    h_search = _dev.empty(n)
    _hip.call("nb_search_lengths", _hip.ptr(h_guess_d), float(h_max), float(skin), float(box_cap), n,
              _hip.ptr(h_search))  # fmt: skip
    if m_uniform is None:
        m_uniform = _dev.uniform_mass(mass_d)
    builds = 0
    grown = np.zeros(1, np.int64)
    for _ in range(12):
        if blocked:
            nb = build_lists_blocked(pos_d, h_search, box_arr, capacity_hint)
        else:
            nb = build_lists(pos_d, h_search, box_arr, capacity_hint)
        builds += 1
        h, rho, omega, it, fl = density_device(pos_d, mass_d, h_guess_d, nb, h_search, kernel_name, eta, tol,
                                               h_max, rows_d, nrows, m_uniform)  # fmt: skip
        _hip.call("nb_grow_search", _hip.ptr(fl), _hip.ptr(h), _hip.ptr(rows_d), nrows, float(box_cap),
                  _hip.ptr(h_search), _hip.ptr(h_guess_d), grown.ctypes.data_as(_hip._P64))  # fmt: skip
        if grown[0] == 0:
            return DensityResult(h, rho, omega, it, fl, nb, h_search if keep_search else None, builds)
        capacity_hint = max(capacity_hint, nb.row_cap if blocked else nb.n_entries)
    raise RuntimeError("smoothing lengths kept outgrowing the neighbour search")


def hydro_density_device(pos_d, mass_d, h_guess_d, kernel_name, eta, tol, h_max, skin, m_uniform,
                         precision="fp64", h_start_d=None, index_buffer=None, row_hint=None,
                         row_counts=None):  # fmt: skip
    """The density pass of the coupled step (open box): compute_density's rebuild rounds
    (sph.py:329-387) on the fused gather + h-rho solve kernel, then the tightened lists at the
    final h, in 64-row blocks with capacities of their own. h_guess_d is overwritten.
    row_hint / row_counts: (N, 2) int32 device tensors, the previous pass's row counts of the
    same particles (sizes the blocks) and this pass's (filled). In FP64, particles stored in grid
    order (the coupled step keeps them so) take the lane-group gather (nb_hydro_density64: the
    same rows, sums and bits), others the per-thread one. Returns (DensityResult with the tight
    lists as ``neighbours``, search-list entries)."""
    from .neighbours import BlockedLists

    n = int(pos_d.shape[0])
    kid = KERNEL_IDS[kernel_name]
    h_search = _dev.empty(n)
    _hip.call("nb_search_lengths", _hip.ptr(h_guess_d), float(h_max), float(skin), float(np.inf), n,
              _hip.ptr(h_search))  # fmt: skip
    h, rho, omega = (_dev.empty(n) for _ in range(3))
    iters = _dev.empty(n, torch.int32)
    flags = _dev.empty(n, torch.int32)
    stats = np.zeros(4, np.int64)
    grown = np.zeros(1, np.int64)
    lanes = False
    for builds in range(1, 13):
        start = h_start_d if (h_start_d is not None and builds == 1) else h_guess_d
        hint = row_hint if builds == 1 or row_counts is None else row_counts
        common = (n, _hip.ptr(hint), _hip.ptr(row_counts), kid, float(C_NORM[kid]), float(eta), float(tol),
                  float(h_max))  # fmt: skip
        outs = (_hip.ptr(h), _hip.ptr(rho), _hip.ptr(omega), _hip.ptr(iters), _hip.ptr(flags),
                stats.ctypes.data_as(_hip._P64))  # fmt: skip
        args = (_hip.ptr(pos_d), _hip.ptr(mass_d), float(m_uniform), _hip.ptr(start), _hip.ptr(h_search))
        lanes = (precision == "fp64" and FP64_LANE_LISTS
                 and _hip.call("nb_hydro_density64", *args, *common, *outs) != NB_NOT_GRID_ORDER)  # fmt: skip
        if not lanes:
            _hip.call("nb_hydro_density", *args, *common, PRECISIONS[precision], *outs)
        _hip.call("nb_grow_search", _hip.ptr(flags), _hip.ptr(h), None, n, float(np.inf), _hip.ptr(h_search),
                  _hip.ptr(h_guess_d), grown.ctypes.data_as(_hip._P64))  # fmt: skip
        if grown[0] == 0:
            break
    else:
        raise RuntimeError("smoothing lengths kept outgrowing the neighbour search")
    entries = max(int(stats[3]), 1)
    ln = _dev.empty(n, torch.int32)
    boff = _dev.empty((n + 63) // 64 + 1, torch.int64)
    idx = _dev.empty(entries, torch.int32) if index_buffer is None else index_buffer.take(entries)
    total = np.zeros(1, np.int64)
    if lanes:
        _hip.call("nb_hydro_tight64", _hip.ptr(pos_d), _hip.ptr(h), _hip.ptr(boff), _hip.ptr(ln), _hip.ptr(idx),
                  entries, total.ctypes.data_as(_hip._P64))  # fmt: skip
    else:
        _hip.call("nb_hydro_tight", _hip.ptr(h), _hip.ptr(boff), _hip.ptr(ln), _hip.ptr(idx), entries,
                  total.ctypes.data_as(_hip._P64))  # fmt: skip
    tight = BlockedLists(ln, idx, 0, int(total[0]), None, block_offset=boff)
    return DensityResult(h, rho, omega, iters, flags, tight, None, builds), int(stats[2])


def hydro_density_mixed_device(pos_d, mass_d, h_guess_d, kernel_name, eta, tol, h_max, skin, m_uniform,
                               h_start_d=None, cloud_d=None, seg_off=None, cell_factor=1.0,
                               index_buffer=None, row_hint=None, row_counts=None):  # fmt: skip
    """hydro_density_device in mixed precision: the same exact search lists, FP32 compensated
    sums in the h-rho solve (Newton steps and convergence test in FP64), tight lists in rows of
    their own capacity. row_hint / row_counts: (N, 2) int32 device tensors, the previous pass's
    row counts of the same particles (sizes the rows) and this pass's (filled). Returns
    (DensityResult with RowLists as ``neighbours``, search-list entries)."""
    from .neighbours import RowLists

    n = int(pos_d.shape[0])
    kid = KERNEL_IDS[kernel_name]
    off = None if seg_off is None else np.ascontiguousarray(seg_off, np.int64)
    h_search = _dev.empty(n)
    _hip.call("nb_search_lengths", _hip.ptr(h_guess_d), float(h_max), float(skin), float(np.inf), n,
              _hip.ptr(h_search))  # fmt: skip
    h, rho, omega = (_dev.empty(n) for _ in range(3))
    iters = _dev.empty(n, torch.int32)
    flags = _dev.empty(n, torch.int32)
    stats = np.zeros(4, np.int64)
    grown = np.zeros(1, np.int64)
    for builds in range(1, 13):
        start = h_start_d if (h_start_d is not None and builds == 1) else h_guess_d
        _hip.call("nb_hydro_density_mixed", _hip.ptr(pos_d), _hip.ptr(mass_d), float(m_uniform), _hip.ptr(start),
                  _hip.ptr(h_search), n, _hip.ptr(cloud_d), None if off is None else off.ctypes.data_as(_hip._P64),
                  0 if off is None else len(off) - 1, float(cell_factor),
                  _hip.ptr(row_hint if builds == 1 or row_counts is None else row_counts), _hip.ptr(row_counts),
                  kid, float(C_NORM[kid]), float(eta),
                  float(tol), float(h_max), _hip.ptr(h),
                  _hip.ptr(rho), _hip.ptr(omega), _hip.ptr(iters), _hip.ptr(flags),
                  stats.ctypes.data_as(_hip._P64))  # fmt: skip
        _hip.call("nb_grow_search", _hip.ptr(flags), _hip.ptr(h), None, n, float(np.inf), _hip.ptr(h_search),
                  _hip.ptr(h_guess_d), grown.ctypes.data_as(_hip._P64))  # fmt: skip
        if grown[0] == 0:
            break
    else:
        raise RuntimeError("smoothing lengths kept outgrowing the neighbour search")
    entries = int(stats[2])
    ln = _dev.empty(n, torch.int32)
    off_t = _dev.empty(n + 1, torch.int64)
    idx = _dev.empty(entries, torch.int32) if index_buffer is None else index_buffer.take(entries)
    total = np.zeros(1, np.int64)
    _hip.call("nb_hydro_tight_mixed", _hip.ptr(pos_d), _hip.ptr(h), _hip.ptr(rho), _hip.ptr(omega),
              _hip.ptr(off_t), _hip.ptr(ln), _hip.ptr(idx), entries, total.ctypes.data_as(_hip._P64))  # fmt: skip
    tight = RowLists(ln, idx, off_t, int(total[0]), None)
    return DensityResult(h, rho, omega, iters, flags, tight, None, builds), int(stats[2])


# --------------------------------------------------------------------------- forces


@dataclass(frozen=True)
class ShockCapturing:
    """Shock-capturing settings (sph.py:409-429)."""

    beta: float = 2.0
    alpha_u: float = 1.0
    conductivity_signal: int = COND_PRESSURE
    reconstruction: bool = False
    closeness: bool = True
    eta_crit: float = 1.0 / DEFAULT_ETA
    eta_fold: float = 0.2
    recon_gate: int = GATE_MINMOD


DEFAULT_SHOCK = ShockCapturing()


@dataclass
class ForceResult:
    acc: object
    dudt: object
    vsig_max: object
    dudt_visc: object = None
    dudt_cond: object = None


def force_params(kernel_name, use_visc, use_cond, shock: ShockCapturing, precision="fp64") -> _hip.ForceParams:
    kid = KERNEL_IDS[kernel_name]
    p = _hip.ForceParams()
    p.kid = kid
    p.use_visc = int(use_visc)
    p.use_cond = int(use_cond)
    p.cond_signal = int(shock.conductivity_signal)
    p.use_recon = int(use_visc and shock.reconstruction)
    p.closeness = int(shock.closeness)
    p.gate = int(shock.recon_gate)
    p.precision = PRECISIONS[precision]
    p.cnorm = float(C_NORM[kid])
    p.beta = float(shock.beta)
    p.alpha_u = float(shock.alpha_u)
    p.eta_crit = float(shock.eta_crit)
    p.eta_fold = float(shock.eta_fold)
    return p


def forces_device(pos_d, vel_d, mass_d, h_d, rho_d, omega_d, p_d, nb, kernel_name, u_d, cs_d, alpha_d, shock,
                  rows_d=None, nrows=None, gv_d=None, precision="fp64"):  # fmt: skip
    n = int(pos_d.shape[0])
    nrows = n if nrows is None else nrows
    use_visc = alpha_d is not None and cs_d is not None
    use_cond = use_visc and u_d is not None
    prm = force_params(kernel_name, use_visc, use_cond, shock, precision)
    acc = _dev.empty((nrows, 3))
    dudt, vsig, vheat, cheat = (_dev.empty(nrows) for _ in range(4))
    _hip.call("nb_forces", _hip.ptr(pos_d), _hip.ptr(vel_d), _hip.ptr(mass_d), _hip.ptr(h_d), _hip.ptr(rho_d),
              _hip.ptr(omega_d), _hip.ptr(p_d), _hip.ptr(u_d if use_cond else None),
              _hip.ptr(cs_d if use_visc else None), _hip.ptr(alpha_d if use_visc else None),
              _hip.ptr(gv_d if prm.use_recon else None), ctypes.byref(nb.desc()),
              _hip.box_arg(nb.box), _hip.ptr(rows_d), nrows, prm, _hip.ptr(acc), _hip.ptr(dudt),
              _hip.ptr(vsig), _hip.ptr(vheat), _hip.ptr(cheat))  # fmt: skip
    return ForceResult(acc, dudt, vsig, vheat, cheat)


def forces_numba(pos, vel, mass, h, rho, omega, pressure, nb: Neighbours, kernel_name=DEFAULT_KERNEL, u=None,
                 c_s=None, alpha=None, shock: ShockCapturing = DEFAULT_SHOCK, rows=None, grad_v=None,
                 precision="fp64") -> ForceResult:  # fmt: skip
    """Force pass (sph.py:655-690). Viscosity is on when ``alpha`` and ``c_s`` are given;
    conductivity when ``u`` is given as well; ``grad_v`` (N, 3, 3) with reconstruction."""
    as_dev = _dev.is_dev(pos)
    n = len(pos)
    use_visc = alpha is not None and c_s is not None
    if use_visc and shock.reconstruction and (grad_v is None or tuple(np.shape(grad_v)) != (n, 3, 3)):
        raise ValueError("velocity reconstruction needs grad_v of shape (N, 3, 3)")
    pos_d = _dev.f64(pos)
    rows_d, nrows = _dev.rows_dev(rows, n)

    def opt(a):
        return None if a is None else _dev.f64(a)

    f = forces_device(pos_d, _dev.f64(vel), _dev.f64(mass), _dev.f64(h), _dev.f64(rho), _dev.f64(omega),
                      _dev.f64(pressure), nb, kernel_name, opt(u), opt(c_s), opt(alpha), shock, rows_d, nrows,
                      opt(grad_v) if (use_visc and shock.reconstruction) else None, precision)  # fmt: skip
    if as_dev:
        return f
    return ForceResult(*(_dev.host(t) for t in (f.acc, f.dudt, f.vsig_max, f.dudt_visc, f.dudt_cond)))


# --------------------------------------------------------------------------- gradients


@dataclass
class Gradients:
    div_v: object
    curl_v: object
    ddivv_dt: object


SINGULAR_TOL = 1e-6


def gradients_device(pos_d, vel_d, acc_d, mass_d, h_d, nb, kernel_name, rows_d=None, nrows=None,
                     precision="fp64") -> Gradients:  # fmt: skip
    n = int(pos_d.shape[0])
    nrows = n if nrows is None else nrows
    kid = KERNEL_IDS[kernel_name]
    div, curl, ddiv = (_dev.empty(nrows) for _ in range(3))
    entry = "nb_gradients_mixed" if PRECISIONS[precision] == 1 else "nb_gradients"
    _hip.call(entry, _hip.ptr(pos_d), _hip.ptr(vel_d), _hip.ptr(acc_d), _hip.ptr(mass_d), _hip.ptr(h_d),
              ctypes.byref(nb.desc()), _hip.box_arg(nb.box), _hip.ptr(rows_d), nrows, kid,
              float(C_NORM[kid]), _hip.ptr(div), _hip.ptr(curl), _hip.ptr(ddiv))  # fmt: skip
    return Gradients(div, curl, ddiv)


def gradients_numba(pos, vel, acc, mass, h, nb: Neighbours, kernel_name=DEFAULT_KERNEL, rows=None,
                    precision="fp64") -> Gradients:  # fmt: skip
    """Corrected velocity / acceleration gradients (sph.py:887-902): div v, |curl v|,
    d(div v)/dt. ``acc`` None means zero accelerations. ``precision`` "mixed": FP32 pair
    terms with compensated sums."""
    as_dev = _dev.is_dev(pos)
    pos_d = _dev.f64(pos)
    rows_d, nrows = _dev.rows_dev(rows, int(pos_d.shape[0]))
    g = gradients_device(pos_d, _dev.f64(vel), None if acc is None else _dev.f64(acc), _dev.f64(mass),
                         _dev.f64(h), nb, kernel_name, rows_d, nrows, precision)  # fmt: skip
    return g if as_dev else Gradients(_dev.host(g.div_v), _dev.host(g.curl_v), _dev.host(g.ddivv_dt))


def velocity_gradients_numba(pos, vel, mass, h, nb: Neighbours, kernel_name=DEFAULT_KERNEL, rows=None, out=None):
    """Corrected velocity gradient G_ij = dv^i/dx^j for ``rows`` (sph.py:818-834); fills
    ``out`` (N, 3, 3; zeros if not given) and returns it."""
    as_dev = _dev.is_dev(pos)
    pos_d, vel_d, mass_d, h_d = _dev.f64(pos), _dev.f64(vel), _dev.f64(mass), _dev.f64(h)
    n = int(pos_d.shape[0])
    rows_d, nrows = _dev.rows_dev(rows, n)
    out_d = torch.zeros((n, 3, 3), dtype=torch.float64, device=pos_d.device) if out is None else _dev.f64(out)
    velocity_gradients_device(pos_d, vel_d, mass_d, h_d, nb, kernel_name, out_d, rows_d, nrows)
    if as_dev:
        return out_d
    res = _dev.host(out_d)
    if out is not None:
        out[...] = res
        return out
    return res


def velocity_gradients_device(pos_d, vel_d, mass_d, h_d, nb, kernel_name, out_d, rows_d=None, nrows=None):
    nrows = int(pos_d.shape[0]) if nrows is None else nrows
    kid = KERNEL_IDS[kernel_name]
    _hip.call("nb_velocity_gradients", _hip.ptr(pos_d), _hip.ptr(vel_d), _hip.ptr(mass_d), _hip.ptr(h_d),
              ctypes.byref(nb.desc()), _hip.box_arg(nb.box), _hip.ptr(rows_d), nrows, kid,
              float(C_NORM[kid]), _hip.ptr(out_d))  # fmt: skip
    return out_d


def ns_gradients_numba(pos, vel, mass, h, rho, omega, nb: Neighbours, kernel_name=DEFAULT_KERNEL, rows=None,
                       out=None):  # fmt: skip
    """Velocity gradient with the SPH difference operator (sph.py:984-999), (N, 3, 3)."""
    as_dev = _dev.is_dev(pos)
    pos_d, vel_d, mass_d, h_d, rho_d, om_d = (_dev.f64(a) for a in (pos, vel, mass, h, rho, omega))
    n = int(pos_d.shape[0])
    rows_d, nrows = _dev.rows_dev(rows, n)
    out_d = torch.zeros((n, 3, 3), dtype=torch.float64, device=pos_d.device) if out is None else _dev.f64(out)
    ns_gradients_device(pos_d, vel_d, mass_d, h_d, rho_d, om_d, nb, kernel_name, out_d, rows_d, nrows)
    if as_dev:
        return out_d
    res = _dev.host(out_d)
    if out is not None:
        out[...] = res
        return out
    return res


def ns_gradients_device(pos_d, vel_d, mass_d, h_d, rho_d, om_d, nb, kernel_name, out_d, rows_d=None, nrows=None):
    nrows = int(pos_d.shape[0]) if nrows is None else nrows
    kid = KERNEL_IDS[kernel_name]
    _hip.call("nb_ns_gradients", _hip.ptr(pos_d), _hip.ptr(vel_d), _hip.ptr(mass_d), _hip.ptr(h_d),
              _hip.ptr(rho_d), _hip.ptr(om_d), ctypes.byref(nb.desc()), _hip.box_arg(nb.box),
              _hip.ptr(rows_d), nrows, kid, float(C_NORM[kid]), _hip.ptr(out_d))  # fmt: skip
    return out_d


def ns_forces_numba(pos, mass, h, rho, omega, stress, nb: Neighbours, kernel_name=DEFAULT_KERNEL, rows=None):
    """Viscous acceleration of the Navier-Stokes stress (sph.py:1030-1045), (len(rows), 3)."""
    as_dev = _dev.is_dev(pos)
    pos_d, mass_d, h_d, rho_d, om_d, st_d = (_dev.f64(a) for a in (pos, mass, h, rho, omega, stress))
    rows_d, nrows = _dev.rows_dev(rows, int(pos_d.shape[0]))
    acc = ns_forces_device(pos_d, mass_d, h_d, rho_d, om_d, st_d, nb, kernel_name, rows_d, nrows)
    return acc if as_dev else _dev.host(acc)


def ns_forces_device(pos_d, mass_d, h_d, rho_d, om_d, st_d, nb, kernel_name, rows_d=None, nrows=None):
    nrows = int(pos_d.shape[0]) if nrows is None else nrows
    acc = _dev.empty((nrows, 3))
    kid = KERNEL_IDS[kernel_name]
    _hip.call("nb_ns_forces", _hip.ptr(pos_d), _hip.ptr(mass_d), _hip.ptr(h_d), _hip.ptr(rho_d), _hip.ptr(om_d),
              _hip.ptr(st_d), ctypes.byref(nb.desc()), _hip.box_arg(nb.box), _hip.ptr(rows_d),
              nrows, kid, float(C_NORM[kid]), _hip.ptr(acc))  # fmt: skip
    return acc


def disc_forces_numba(pos, vel, mass, h, rho, omega, c_s, alpha_av, nb: Neighbours, kernel_name=DEFAULT_KERNEL,
                      rows=None):  # fmt: skip
    """Disc viscosity through the AV form on every pair (sph.py:1120-1137): (acc, du/dt)."""
    as_dev = _dev.is_dev(pos)
    pos_d, vel_d, mass_d, h_d, rho_d, om_d, cs_d = (_dev.f64(a) for a in (pos, vel, mass, h, rho, omega, c_s))
    rows_d, nrows = _dev.rows_dev(rows, int(pos_d.shape[0]))
    acc, dudt = disc_forces_device(pos_d, vel_d, mass_d, h_d, rho_d, om_d, cs_d, alpha_av, nb, kernel_name,
                                   rows_d, nrows)  # fmt: skip
    return (acc, dudt) if as_dev else (_dev.host(acc), _dev.host(dudt))


def disc_forces_device(pos_d, vel_d, mass_d, h_d, rho_d, om_d, cs_d, alpha_av, nb, kernel_name, rows_d=None,
                       nrows=None):  # fmt: skip
    nrows = int(pos_d.shape[0]) if nrows is None else nrows
    acc = _dev.empty((nrows, 3))
    dudt = _dev.empty(nrows)
    kid = KERNEL_IDS[kernel_name]
    _hip.call("nb_disc_forces", _hip.ptr(pos_d), _hip.ptr(vel_d), _hip.ptr(mass_d), _hip.ptr(h_d),
              _hip.ptr(rho_d), _hip.ptr(om_d), _hip.ptr(cs_d), float(alpha_av), ctypes.byref(nb.desc()),
              _hip.box_arg(nb.box), _hip.ptr(rows_d), nrows, kid, float(C_NORM[kid]),
              _hip.ptr(acc), _hip.ptr(dudt))  # fmt: skip
    return acc, dudt
