"""Run folders, manifests, checkpoints and snapshots of device-resident runs (runio.py of the
reference, runio.py:1-100).

The call surface is the reference's: ``new_run``, ``update_manifest``, ``save_checkpoint``,
``load_checkpoint``, ``save_snapshot``. A run lives in ``<runs>/<name>/`` (``runs`` is
``$NEBULA_RUNS`` or ``./runs``) and is never overwritten; ``manifest.json`` records the git commit
of the working directory (and whether it had uncommitted changes), the configuration, library
versions, the machine and the GPU, and start / end / resume times. ``checkpoint.pkl`` holds the
whole float64 payload and is written atomically (a temporary file renamed over the old one), so
a crash never leaves half a checkpoint; snapshots for plots and rendering are float32 ``.npz``.

Device state in a payload: a ``GasState`` pickles its GPU tensors as host arrays (bit for bit,
in its stored particle order, with ``perm`` mapping back to the caller's) and unpickles onto the
current GPU, so a run resumed from a checkpoint continues exactly as the uninterrupted run
would. Scratch that a step rebuilds anyway (the tight-list buffer) is not stored.
``snapshot_arrays`` gives per-particle fields in the caller's order for ``save_snapshot``.
"""

from __future__ import annotations

import json
import os
import pickle
import platform
import subprocess
import time
from dataclasses import asdict, is_dataclass
from pathlib import Path

import numpy as np


def runs_root() -> Path:
    return Path(os.environ.get("NEBULA_RUNS", Path.cwd() / "runs"))


def _git() -> dict:
    def run(*args):
        r = subprocess.run(["git", *args], cwd=Path.cwd(), capture_output=True, text=True, check=False)
        return r.stdout.strip()

    return {"commit": run("rev-parse", "HEAD"), "uncommitted_changes": bool(run("status", "--porcelain"))}


def _versions() -> dict:
    import scipy
    import torch

    return {"python": platform.python_version(), "numpy": np.__version__, "scipy": scipy.__version__,
            "torch": torch.__version__, "hip": getattr(torch.version, "hip", None)}  # fmt: skip


def _gpu() -> str | None:
    try:
        import torch

        return torch.cuda.get_device_name(0) if torch.cuda.is_available() else None
    except Exception:
        return None


def _jsonable(obj):
    if is_dataclass(obj) and not isinstance(obj, type):
        return {k: _jsonable(v) for k, v in asdict(obj).items()}
    if isinstance(obj, dict):
        return {k: _jsonable(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_jsonable(v) for v in obj]
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, float) and not np.isfinite(obj):
        return str(obj)
    if isinstance(obj, (str, int, float, bool)) or obj is None:
        return obj
    return repr(obj)  # a model object, a tensor: its description, not its contents


def new_run(name: str, config: dict, runs: Path | None = None) -> Path:
    """Create <runs>/<name>/ with a snapshots/ folder (refuses an existing one) and write its
    manifest (runio.py:58-74)."""
    folder = Path(runs if runs is not None else runs_root()) / name
    folder.mkdir(parents=True, exist_ok=False)
    (folder / "snapshots").mkdir()
    manifest = {
        "name": name, "git": _git(), "config": _jsonable(config), "versions": _versions(),
        "machine": {"platform": platform.platform(), "processor": platform.processor(), "machine": platform.machine(),
                    "gpu": _gpu()},
        "seed": None, "start": time.strftime("%Y-%m-%d %H:%M:%S %Z"), "end": None, "resumed": [],
    }  # fmt: skip
    (folder / "manifest.json").write_text(json.dumps(manifest, indent=2))
    return folder


def update_manifest(folder: Path, **fields) -> None:
    """Set manifest fields; ``resumed=<time>`` appends a resume record (runio.py:77-85)."""
    path = Path(folder) / "manifest.json"
    manifest = json.loads(path.read_text())
    for k, v in fields.items():
        if k == "resumed":
            manifest["resumed"].append({"time": v, "git": _git()})
        else:
            manifest[k] = _jsonable(v)
    path.write_text(json.dumps(manifest, indent=2))


def save_checkpoint(folder: Path, payload) -> None:
    """Pickle ``payload`` (the full float64 state, GasStates included) atomically
    (runio.py:85-89)."""
    folder = Path(folder)
    tmp = folder / "checkpoint.tmp"
    tmp.write_bytes(pickle.dumps(payload, protocol=pickle.HIGHEST_PROTOCOL))
    tmp.replace(folder / "checkpoint.pkl")


def load_checkpoint(folder: Path):
    """The payload of a checkpoint this code wrote (runio.py:92-93); GasStates come back on the
    current GPU."""
    return pickle.loads((Path(folder) / "checkpoint.pkl").read_bytes())


def save_snapshot(folder: Path, index: int, t: float, **arrays) -> None:
    """float32 copies of the given per-particle arrays (NumPy or device tensors), integers kept,
    in snapshots/snap_<index>.npz with the time (runio.py:96-100)."""
    out = {}
    for k, v in arrays.items():
        a = v.detach().cpu().numpy() if hasattr(v, "detach") else np.asarray(v)
        out[k] = a.astype(np.float32) if np.issubdtype(a.dtype, np.floating) else a
    np.savez_compressed(Path(folder) / "snapshots" / f"snap_{index:05d}.npz", t=np.float64(t), **out)


def snapshot_arrays(st, *names) -> dict:
    """Fields of a GasState as host arrays in the caller's particle order (GasState keeps its
    particles in the neighbour grid's order; ``perm`` maps back)."""
    host = st.to_numpy()
    return {k: host[k] for k in names}
