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


def test_product_never_imports_the_oracle():
    pkg = REPO / "blackhole-nebula-simulator_amd"
    offenders = [str(p) for p in pkg.rglob("*.py") if re.search(r"^\s*(from|import)\s+oracle", p.read_text(), re.M)]
    check("no oracle import in the product package", [], offenders)
