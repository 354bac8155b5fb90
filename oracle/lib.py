"""ctypes bindings of the oracle's C library (test infrastructure only).

The library is built by ``make -C oracle`` (``oracle/_build/libnebula_oracle.so``).
"""

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "libnebula_oracle.so"


def build() -> Path:
    """Compile the oracle library (gcc, OpenMP); returns its path."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        build()
    return ctypes.CDLL(str(LIB_PATH))


_lib = _load()

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


def _sig(name, *args, res=None):
    fn = getattr(_lib, name)
    fn.argtypes = list(args)
    fn.restype = res
    return fn


kernel_f_c = _sig("orc_kernel_f", _f64, _i64, res=_f64)
kernel_df_c = _sig("orc_kernel_df", _f64, _i64, res=_f64)
build_grid = _sig("orc_build_grid", _D, _i64, _D, _D, _I, _I, _I)
gather = _sig("orc_gather", _D, _D, _i64, _D, _D, _D, _I, _I, _I, _I, _I, _I, _i64)
sort_rows = _sig("orc_sort_rows", _I, _i64, _I)
transpose = _sig("orc_transpose", _I, _I, _i64, _I, _I)
merge_rows = _sig("orc_merge_rows", _I, _I, _I, _I, _i64, _I, _I, _I, _i64)
tighten = _sig("orc_tighten", _D, _D, _D, _I, _I, _I, _i64, _I, _I, _I, _i64)
density = _sig("orc_density", _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _f64, _f64, _f64, _I,
               _i64, _D, _D, _D, _I, _I)  # fmt: skip
forces = _sig(
    "orc_forces", _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _i64, _i64,
    _f64, _f64, _i64, _I, _i64, _i64, _i64, _f64, _f64, _i64, _D, _D, _D, _D, _D, _D,
)
gradients = _sig("orc_gradients", _D, _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _I, _i64, _D, _D,
                 _D)  # fmt: skip
velocity_gradients = _sig("orc_velocity_gradients", _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _I,
                          _i64, _D)  # fmt: skip
ns_gradients = _sig("orc_ns_gradients", _D, _D, _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _I, _i64,
                    _D)  # fmt: skip
ns_forces = _sig("orc_ns_forces", _D, _D, _D, _D, _D, _D, _I, _I, _D, _i64, _f64, _I, _i64, _D)
disc_forces = _sig("orc_disc_forces", _D, _D, _D, _D, _D, _D, _D, _f64, _I, _I, _D, _i64, _f64,
                   _I, _i64, _D, _D)  # fmt: skip
glf_step_many = _sig("orc_glf_step_many", _D, _D, _D, _i64, _D, _D, _I)
gn_v_from_p = _sig("orc_gn_v_from_p", _D, _D, _i64, _D)
inv3_det3 = _sig("orc_inv3_det3", _D, _i64, _D, _D)


def d(a: np.ndarray):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def i(a: np.ndarray):
    """Pointer to a C-contiguous int64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def set_threads(n: int) -> None:
    """OpenMP thread count for the following calls (the library reads OMP_NUM_THREADS at
    first use; omp_set_num_threads changes it afterwards)."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    try:
        omp = ctypes.CDLL("libgomp.so.1")
        omp.omp_set_num_threads(ctypes.c_int(n))
    except OSError:
        pass
