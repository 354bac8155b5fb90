"""Run folders, manifests, checkpoints and snapshots (CPU; runio.py:58-100 of the reference)."""

import json

import numpy as np
import pytest

from helpers import check


def test_new_run_manifest_and_refusal(tmp_path):
    from nebula.sim import runio

    folder = runio.new_run("r1", {"n": 4096, "eta": 1.2, "tend": float("inf")}, runs=tmp_path)
    check("snapshots folder", True, (folder / "snapshots").is_dir())
    m = json.loads((folder / "manifest.json").read_text())
    check("config recorded (non-finite floats as text)", {"n": 4096, "eta": 1.2, "tend": "inf"}, m["config"])
    check("git and versions recorded", True, "commit" in m["git"] and "numpy" in m["versions"])
    check("not resumed yet", [], m["resumed"])
    with pytest.raises(FileExistsError):
        runio.new_run("r1", {}, runs=tmp_path)
    runio.update_manifest(folder, resumed="2026-12-04 10:00:00", end="later")
    m = json.loads((folder / "manifest.json").read_text())
    check("resume appended", "2026-12-04 10:00:00", m["resumed"][0]["time"])
    check("field set", "later", m["end"])


def test_checkpoint_round_trip_is_atomic(tmp_path):
    from nebula.sim import runio

    folder = runio.new_run("r2", {}, runs=tmp_path)
    payload = {"x": np.arange(12, dtype=np.float64).reshape(4, 3) / 7.0, "snap": 3, "t": 12.5}
    runio.save_checkpoint(folder, payload)
    check("no temporary file left", False, (folder / "checkpoint.tmp").exists())
    back = runio.load_checkpoint(folder)
    check("float64 arrays come back bit for bit", True, bool(np.array_equal(payload["x"], back["x"])))
    check("scalars", (3, 12.5), (back["snap"], back["t"]))
    runio.save_checkpoint(folder, {"snap": 4})
    check("a new checkpoint replaces the old", 4, runio.load_checkpoint(folder)["snap"])


def test_snapshot_is_float32_with_integers_kept(tmp_path):
    from nebula.sim import runio

    folder = runio.new_run("r3", {}, runs=tmp_path)
    x = np.random.default_rng(1).normal(size=(5, 3))
    runio.save_snapshot(folder, 7, 1.25, x=x, ids=np.arange(5, dtype=np.int64))
    d = np.load(folder / "snapshots" / "snap_00007.npz")
    check("positions stored as float32", np.dtype(np.float32), d["x"].dtype)
    check("values rounded to float32", x.astype(np.float32), d["x"])
    check("integer ids kept", np.dtype(np.int64), d["ids"].dtype)
    check("time stored in float64", 1.25, float(d["t"]))
