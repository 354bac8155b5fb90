"""Writes tests/golden/phantom_force_sample.npz from the reference's own fixture.

The reference holds one Phantom derivative evaluation (tests/data/phantom_force_sample.npz of
the reference repository, written by its scripts/phantom_force_check.py; GPL-3.0 data): 8,931
particles within 1,800 of the hole, Wendland C4 kernel, gas + radiation pressure from Phantom,
viscosity alpha per particle, velocity-signal conductivity (alpha_u 0.05), h_max 200, and
Phantom's hydro acceleration and du/dt with velocity reconstruction on and off. Its own test,
tests/sim/test_phantom_patch.py:37-107, compares the particles within 600 to a median relative
difference of 1e-5 and a 99th percentile of 1e-3.

This script copies the arrays unchanged (a plain data load, no pickle) into a compressed file
beside it, so that the tests can use them where the reference is not present.

    python tests/golden/make_phantom_fixture.py
"""

from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tests/data/phantom_force_sample.npz")
DST = Path(__file__).resolve().parent / "phantom_force_sample.npz"


def main() -> None:
    with np.load(SRC, allow_pickle=False) as d:
        arrays = {k: np.array(d[k]) for k in d.files}
    np.savez_compressed(DST, **arrays)
    print(f"{DST}: {', '.join(f'{k} {v.shape}' for k, v in arrays.items())}, {DST.stat().st_size} bytes")


if __name__ == "__main__":
    main()
