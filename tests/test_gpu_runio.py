"""Checkpoint and resume of device-resident runs (GPU): a run checkpointed after some steps and
resumed from the file continues bit for bit as the uninterrupted run, in FP64 (coupled.step) and
in mixed precision (the single-synchronisation stepper); snapshots are written in the caller's
particle order."""

import numpy as np
import pytest

from conftest import gpu_available
from helpers import bit_equal, check

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FIELDS = ("x", "p", "u", "h", "rho", "alpha", "a_hydro", "dudt", "vsig_max", "c_s", "omega")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resume_from_checkpoint_is_bit_identical(tmp_path, precision):
    from dataclasses import replace

    from nebula.ics import stream_crash
    from nebula.sim import coupled, runio
    from nebula.sim.stepper import Stepper

    ic = stream_crash(8192, length=200.0, seed=20261014)
    cfg = replace(coupled.baseline_config(balsara=True, precision=precision), cooling_powerlaw=True)
    print(f"checkpoint: 8192-particle stream crash, {precision}, 3 + 4 steps", flush=True)

    def advance(st, k):
        if precision == "mixed":
            sp = Stepper(st, cfg)
            for _ in range(k):
                sp.step()
            return sp.state
        for _ in range(k):
            st = coupled.step(st, cfg, coupled.stable_dt(st, cfg))
        return st

    st = advance(coupled.initialise(ic.x, ic.v, ic.u, ic.mass, ic.h, cfg), 3)
    folder = runio.new_run(f"resume_{precision}", {"n": len(ic.x), "precision": precision}, runs=tmp_path)
    runio.save_checkpoint(folder, {"st": st, "snap": 1})
    snap = runio.snapshot_arrays(st, "x", "u", "rho")
    runio.save_snapshot(folder, 1, st.t, **snap)
    straight = advance(st, 4)
    back = runio.load_checkpoint(folder)
    check("payload fields", 1, back["snap"])
    resumed = advance(back["st"], 4)
    a, b = straight.to_numpy(), resumed.to_numpy()
    for f in FIELDS:
        bit_equal(f"{precision}: {f} after resume", a[f], b[f])
    check("time", straight.t, resumed.t)
    check("radiated energy", straight.e_radiated, resumed.e_radiated)
    d = np.load(folder / "snapshots" / "snap_00001.npz")
    check("snapshot in the caller's order (float32 of the state)", snap["x"].astype(np.float32), d["x"])
