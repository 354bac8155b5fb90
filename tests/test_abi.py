"""The C ABI (CPU only, no compute calls): libnebula_hip.so loads, exports every entry point
include/nebula_hip.h declares, and the Python binding declares exactly those; the product
package has no path to the oracle."""

import re
from pathlib import Path

from helpers import check

REPO = Path(__file__).resolve().parents[1]


def declared():
    text = (REPO / "include" / "nebula_hip.h").read_text()
    return sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from nebula import _hip

    lib = _hip.load_library()
    names = declared()
    missing = [n for n in names if not hasattr(lib, n)]
    check("header declares entry points", True, len(names) > 20)
    check("every declared entry point exported", [], missing)
    check("binding declares exactly the header's entry points", names, sorted(_hip.SIGNATURES))
    check("ABI version", 4, lib.nb_abi_version())


def declared_arities():
    """Parameter count of every prototype in the header (comments stripped)."""
    text = (REPO / "include" / "nebula_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(nb_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_binding_arities_match_the_header():
    """Each ctypes signature takes as many arguments as the C prototype (a missing pointer would
    shift every later argument)."""
    from nebula import _hip

    ar = declared_arities()
    wrong = {n: (len(_hip.SIGNATURES[n][1]), k) for n, k in ar.items() if len(_hip.SIGNATURES[n][1]) != k}
    check("prototypes parsed", True, len(ar) > 20)
    check("ctypes argument counts = header parameter counts", {}, wrong)


def test_product_never_imports_the_oracle():
    pkg = REPO / "blackhole-nebula-simulator_amd"
    offenders = [str(p) for p in pkg.rglob("*.py") if re.search(r"^\s*(from|import)\s+oracle", p.read_text(), re.M)]
    check("no oracle import in the product package", [], offenders)
