"""Pinning the oracle's reuse of cached neighbour lists (sph.py:62-147 and 329-387; CPU only).

The reuse rule is a proof: whenever reuse_limits returns h_eff, the cached lists hold every pair
with r <= 2 max(h_lim_i, h_lim_j) at the current positions. These tests restate the reference's
tests/sim/test_neighbour_reuse.py and check that claim by brute force over all pairs, show that
the brute-force check catches a rule with a dropped term (each particle's own move), and that a
density pass on cached lists agrees with one on freshly built lists to the Newton tolerance (a
capped particle's density bit for bit, as in the reference).
"""

import numpy as np

from helpers import check

from oracle import sph as osph
from oracle.neighbours import find_neighbours_grid


def missing_pairs(nb, pos, h_lim):
    """Pairs with r <= 2 max(h_lim_i, h_lim_j) at pos that are not in the lists."""
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    need = d <= 2.0 * np.maximum(h_lim[:, None], h_lim[None, :])
    have = np.zeros_like(need)
    for i in range(len(pos)):
        have[i, nb.of(i)] = True
    return int(np.sum(need & ~have))


def reuse_trial(rng, n, trial):
    """One case of test_neighbour_reuse.py:22-59: a cluster whose particles moved a little, and
    a few (the halo, or particles inside the cluster) that moved far. Returns (cache, pos, h_guess,
    rows, others)."""
    pos0 = rng.normal(0.0, 4.0, (n, 3))
    h = rng.uniform(0.6, 1.4, n) * (1.0 + 0.3 * np.linalg.norm(pos0, axis=1) / 4.0)
    n_fast = trial % 6
    fast = rng.choice(n, n_fast, replace=False)
    halo = trial % 2 == 0
    if halo:
        u = rng.normal(0.0, 1.0, (n_fast, 3))
        pos0[fast] = 18.0 * u / np.linalg.norm(u, axis=1)[:, None]
        h[fast] = rng.uniform(3.0, 6.0, n_fast)
    h_search = 1.4 * h
    cache = osph.NeighbourCache(find_neighbours_grid(pos0, h_search), pos0.copy(), h_search.copy())
    scale = [0.005, 0.02, 0.05, 0.1][trial % 4]
    pos = pos0 + rng.normal(0.0, scale, (n, 3))
    pos[fast] += rng.normal(0.0, 0.6 if halo else 1.5, (n_fast, 3))
    rows = np.sort(rng.choice(n, n // 5, replace=False))
    others = np.ones(n, bool)
    others[rows] = False
    h_guess = h * rng.uniform(0.97, 1.03, n)
    return cache, pos, h_guess, rows, others


def test_reused_lists_hold_every_interacting_pair(rng):
    reused = {"none fast": 0, "with fast set": 0}
    missing = 0
    for trial in range(120):
        cache, pos, h_guess, rows, others = reuse_trial(rng, 800, trial)
        h_eff = osph.reuse_limits(cache, pos, h_guess, rows, others, None)
        if h_eff is None:
            continue
        reused["none fast" if trial % 6 == 0 else "with fast set"] += 1
        missing += missing_pairs(cache.neighbours, pos, np.where(others, h_guess, h_eff))
    print(f"reuse decisions checked: {reused}")
    check("reused lists: interacting pairs missing (brute force)", 0, missing)
    check("reuse exercised without and with far-moving particles", True,
          reused["none fast"] > 0 and reused["with fast set"] > 0)  # fmt: skip


def test_brute_force_check_catches_a_dropped_term(rng):
    """A rule that forgets each particle's own move (h_eff = hs - d_rest / 2, fast rows still
    checked) loses pairs, and the brute-force count sees it: the check above has teeth."""
    lost = 0
    for trial in range(60):
        cache, pos, h_guess, rows, others = reuse_trial(rng, 800, trial)
        d = np.linalg.norm(pos - cache.x_ref, axis=1)
        order = np.lexsort((np.arange(len(d)), -d))
        nf = 4
        h_eff = cache.h_search - 0.5 * d[order[nf]]
        h_lim = np.where(others, h_guess, h_eff)
        lost += missing_pairs(cache.neighbours, pos, h_lim)
    check("a rule without the own-move term loses pairs", True, lost > 0)


def test_one_far_mover_does_not_force_a_rebuild(rng):
    """test_neighbour_reuse.py:62-81: one particle moves far, the rest barely. The rule with a
    single bound (hs - max move) rebuilds; the fast-set rule reuses, checking that particle."""
    n = 1000
    pos0 = rng.normal(0.0, 5.0, (n, 3))
    h = np.full(n, 1.0)
    cache = osph.NeighbourCache(find_neighbours_grid(pos0, 1.4 * h), pos0.copy(), 1.4 * h)
    pos = pos0 + rng.normal(0.0, 0.01, (n, 3))
    far = int(np.argmax(np.linalg.norm(pos0, axis=1)))
    pos[far] += np.array([1.0, 0.0, 0.0])
    old_rule = bool(np.all(1.4 * h - np.max(np.linalg.norm(pos - pos0, axis=1)) > 0.5 * h) and np.all(0.4 * h > h))
    h_eff = osph.reuse_limits(cache, pos, h, np.arange(n), np.zeros(n, bool), None)
    check("single-bound rule would rebuild", False, old_rule)
    check("fast-set rule reuses", True, h_eff is not None)
    check("and the lists are complete", 0, missing_pairs(cache.neighbours, pos, h_eff))


def test_particle_at_the_h_cap_keeps_its_list(rng):
    """test_neighbour_reuse.py:84-97: an isolated particle at h_max moves a little: the lists are
    reused (the search keeps its skin above the cap) and its density equals a fresh build's."""
    pos = rng.normal(0.0, 1.0, (600, 3))
    pos[0] = [30.0, 0.0, 0.0]
    mass = np.full(len(pos), 1.0 / len(pos))
    h0 = np.full(len(pos), 0.3)
    first = osph.compute_density(pos, mass, h0, h_max=2.0)
    check("isolated particle sits at the cap", 2.0, float(first.h[0]), rtol=0.0)
    moved = pos.copy()
    moved[0] += [0.05, 0.0, 0.0]
    again = osph.compute_density(moved, mass, first.h, h_max=2.0, cache=first.cache, rows=np.array([0]))
    fresh = osph.compute_density(moved, mass, first.h, h_max=2.0, rows=np.array([0]))
    check("small move of the capped particle: no rebuild", first.cache.builds, again.cache.builds)
    check("one reuse counted", 1, again.cache.reuses)
    check("its density equals a fresh build's", float(fresh.rho[0]), float(again.rho[0]), rtol=0.0)


def test_density_on_reused_lists_matches_fresh_lists(rng):
    """A drifting cloud over several passes, lists carried in the cache: every pass (reused or
    rebuilt) agrees with a pass on lists built afresh to the Newton tolerance (1e-4 of h; the
    search length is the solver's upper bracket, so a reused list's h_eff can change its path
    to the root), and both reuses and rebuilds happen."""
    n = 1500
    pos = rng.normal(0.0, 1.0, (n, 3))
    vel = rng.normal(0.0, 1.0, (n, 3))
    mass = np.full(n, 1.0 / n)
    h = np.full(n, 0.2)
    res = osph.compute_density(pos, mass, h, kernel_name="wendland_c4", eta=1.6)
    cache = res.cache
    for k in range(8):
        pos = pos + 0.004 * (k + 1) * vel
        h = res.h
        res = osph.compute_density(pos, mass, h, kernel_name="wendland_c4", eta=1.6, cache=cache)
        fresh = osph.compute_density(pos, mass, h, kernel_name="wendland_c4", eta=1.6)
        for f, tol in (("h", 1e-4), ("rho", 3e-4), ("omega", 1e-3)):
            a, b = getattr(fresh, f), getattr(res, f)
            print(f"pass {k} ({'reused' if res.cache is cache else 'rebuilt'}): {f} max rel. diff "
                  f"{np.max(np.abs(b / a - 1)):.2e}")  # fmt: skip
            check(f"pass {k}: {f} on cached lists vs fresh lists", a, b, rtol=tol)
        cache = res.cache
    print(f"builds {cache.builds}, reuses {cache.reuses}")
    check("lists were reused", True, cache.reuses >= 2)
    check("lists were rebuilt as the moves grew", True, cache.builds >= 2)
