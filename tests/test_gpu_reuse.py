"""Reuse of cached neighbour lists on the GPU (sph.py:62-147 and 329-387) against the oracle.

- The device's reuse decision (nb_reuse_displacements / nb_reuse_limits / nb_reuse_fast_rows)
  equals the oracle's on the reference's 120 reuse trials: the same cases reused, h_eff bit for
  bit, and the reused lists complete by brute force (tests/sim/test_neighbour_reuse.py:22-59).
- One far mover reuses the lists (test_neighbour_reuse.py:62-81); a capped particle keeps its
  list and its density (test_neighbour_reuse.py:84-97).
- compute_density with the cache carried over a drifting cloud (open, and in a periodic box with
  particles wrapping across it): every pass bit-identical to the oracle's compute_density with
  its own cache (h, rho, omega, iterations, flags), with the same counts of builds and reuses.
"""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check
from test_oracle_reuse import missing_pairs, reuse_trial

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _device_cache(ocache):
    from nebula.sim.neighbours import find_neighbours_numba
    from nebula.sim.sph import NeighbourCache

    nb = find_neighbours_numba(ocache.x_ref, ocache.h_search)
    bit_equal("device lists = oracle lists (offsets)", ocache.neighbours.offset, nb.offset)
    bit_equal("device lists = oracle lists (indices)", ocache.neighbours.index, nb.index)
    return NeighbourCache(nb, ocache.x_ref.copy(), ocache.h_search.copy())


def test_reuse_decisions_match_oracle(rng):
    from nebula.sim.sph import _reuse_limits

    from oracle import sph as osph

    print("reuse: 120 trials of 800 particles, device decision vs oracle", flush=True)
    counts = {"reused": 0, "rebuilt": 0}
    missing = 0
    for trial in range(120):
        ocache, pos, h_guess, rows, others = reuse_trial(rng, 800, trial)
        ref = osph.reuse_limits(ocache, pos, h_guess, rows, others, None)
        got = _reuse_limits(_device_cache(ocache), pos, h_guess, rows, others, None)
        check(f"trial {trial}: same decision", ref is None, got is None)
        if ref is None:
            counts["rebuilt"] += 1
            continue
        counts["reused"] += 1
        bit_equal(f"trial {trial}: h_eff", ref, got)
        missing += missing_pairs(ocache.neighbours, pos, np.where(others, h_guess, got))
    print(f"decisions: {counts}")
    check("reused lists: interacting pairs missing (brute force)", 0, missing)
    check("both decisions exercised", True, counts["reused"] > 0 and counts["rebuilt"] > 0)


def test_one_far_mover_reuses(rng):
    from nebula.sim.neighbours import find_neighbours_numba
    from nebula.sim.sph import NeighbourCache, _reuse_limits

    n = 1000
    pos0 = rng.normal(0.0, 5.0, (n, 3))
    h = np.full(n, 1.0)
    cache = NeighbourCache(find_neighbours_numba(pos0, 1.4 * h), pos0.copy(), 1.4 * h)
    pos = pos0 + rng.normal(0.0, 0.01, (n, 3))
    far = int(np.argmax(np.linalg.norm(pos0, axis=1)))
    pos[far] += np.array([1.0, 0.0, 0.0])
    h_eff = _reuse_limits(cache, pos, h, np.arange(n), np.zeros(n, bool), None)
    check("one far mover: lists reused", True, h_eff is not None)
    check("and complete", 0, missing_pairs(cache.neighbours, pos, h_eff))


def test_capped_particle_keeps_its_list(rng):
    from nebula.sim.sph import compute_density

    pos = rng.normal(0.0, 1.0, (600, 3))
    pos[0] = [30.0, 0.0, 0.0]
    mass = np.full(len(pos), 1.0 / len(pos))
    first = compute_density(pos, mass, np.full(len(pos), 0.3), h_max=2.0)
    check("isolated particle at the cap", 2.0, float(first.h[0]), rtol=0.0)
    moved = pos.copy()
    moved[0] += [0.05, 0.0, 0.0]
    again = compute_density(moved, mass, first.h, h_max=2.0, cache=first.cache, rows=np.array([0]))
    fresh = compute_density(moved, mass, first.h, h_max=2.0, rows=np.array([0]))
    check("no rebuild", first.cache.builds, again.cache.builds)
    check("one reuse", 1, again.cache.reuses)
    check("density equals a fresh build's", float(fresh.rho[0]), float(again.rho[0]), rtol=0.0)


@pytest.mark.parametrize("periodic", [False, True], ids=["open", "box"])
def test_cached_density_sequence_matches_oracle(rng, periodic):
    from nebula.sim.sph import compute_density

    from oracle import sph as osph

    n = 3000
    box = 6.0 if periodic else None
    if periodic:
        pos = rng.uniform(0.0, box, (n, 3))
        h = np.full(n, 0.25)
    else:
        pos = rng.normal(0.0, 1.0, (n, 3))
        h = np.full(n, 0.2)
    vel = rng.normal(0.0, 1.0, (n, 3))
    mass = np.full(n, 1.0 / n)
    kw = dict(box=box, kernel_name="wendland_c4", eta=1.6)
    print(f"cached density: {n} particles, {'periodic box 6' if periodic else 'open'}, 10 passes", flush=True)
    ref = osph.compute_density(pos, mass, h, **kw)
    got = compute_density(pos, mass, h, **kw)
    h = ref.h.copy()
    for k in range(10):
        pos = pos + 0.004 * (k + 1) * vel
        if periodic:
            pos = np.mod(pos, box)
        # every third pass solves a quarter of the particles (the others keep their h)
        rows = None if k % 3 else np.sort(rng.choice(n, n // 4, replace=False))
        ref = osph.compute_density(pos, mass, h, rows=rows, cache=ref.cache, **kw)
        got = compute_density(pos, mass, h, rows=rows, cache=got.cache, **kw)
        for f in ("h", "rho", "omega", "iterations", "flags"):
            bit_equal(f"pass {k}: {f}", getattr(ref, f), getattr(got, f))
        check(f"pass {k}: builds", ref.cache.builds, got.cache.builds)
        check(f"pass {k}: reuses", ref.cache.reuses, got.cache.reuses)
        if rows is None:
            h = ref.h.copy()
        else:
            h[rows] = ref.h
    print(f"builds {got.cache.builds}, reuses {got.cache.reuses}")
    check("lists were reused", True, got.cache.reuses >= 2)
    check("lists were rebuilt", True, got.cache.builds >= 2)
