"""Per-particle GN velocity on the GPU (fastphys.py of the reference)."""

from __future__ import annotations

from .. import _dev, _hip


def gn_v_from_p_device(x_d, p_d, out=None):
    n = int(x_d.shape[0])
    out = _dev.empty((n, 3)) if out is None else out
    _hip.call("nb_gn_v_from_p", _hip.ptr(x_d), _hip.ptr(p_d), n, _hip.ptr(out))
    return out


def gn_v_from_p_nb(x, p):
    """v = A p + C (p.x) x, A = (r - 2)/r, C = -2 (r - 2)/r^4 (fastphys.py:35-42)."""
    out = gn_v_from_p_device(_dev.f64(x), _dev.f64(p))
    return out if _dev.is_dev(x) else _dev.host(out)
