"""Regenerate the basis-set data files in this folder from the reference checkout.

The reference ships public basis-set data (PySCF-packaged cc-pVDZ and
cc-pVDZ-JKFIT, Apache-2.0; a Basis Set Exchange pack with def2-SVP and STO-3G).
This script reads those data files and writes them in this package's own
compact format, one file per basis:

    {"name": ..., "source": ..., "license": ...,
     "elements": {"<Z>": [{"l": L, "exponents": [...], "coefficients": [[...], ...]}]}}

Each coefficient row of a general contraction becomes its own shell when a
basis is expanded onto atoms (python/generativeqc/basis.py:300-325).  Decimal
strings are parsed with float(), the same nearest-double rounding the
reference applies.  Only data is read; nothing from the reference is imported
or run.

Usage: python vibeqc_amd/data/convert_basis_data.py /root/reference
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _canonical(path: Path):
    payload = json.loads(path.read_text())
    elements = {}
    for element in payload["elements"]:
        shells = []
        for shell in element["shells"]:
            shells.append(
                {
                    "l": int(shell["angular_momentum"]),
                    "exponents": [float(e) for e in shell["exponents"]],
                    "coefficients": [[float(c) for c in row] for row in shell["coefficients"]],
                }
            )
        elements[str(element["atomic_number"])] = shells
    return payload["name"], payload["provenance"], elements


def _pack(pack: dict, name: str):
    elements = {}
    for z, shells in pack["bases"][name]["elements"].items():
        out = []
        for shell in shells:
            coefs = shell["coefficients"]
            rows = coefs if coefs and isinstance(coefs[0], list) else [coefs]
            out.append(
                {
                    "l": int(shell["angular_momentum"]),
                    "exponents": [float(e) for e in shell["exponents"]],
                    "coefficients": [[float(c) for c in row] for row in rows],
                }
            )
        elements[str(int(z))] = out
    return elements


def main(reference: str) -> None:
    root = Path(reference)
    identity = root / "benchmarks/results/issue206-practical-auxiliary/identity"
    for stem in ("cc-pvdz", "cc-pvdz-jkfit"):
        name, provenance, elements = _canonical(identity / f"{stem}.json")
        record = {
            "name": name,
            "source": provenance.get("source", ""),
            "license": provenance.get("license", ""),
            "elements": elements,
        }
        (HERE / f"{stem}.json").write_text(json.dumps(record, indent=1, sort_keys=True) + "\n")
    pack = json.loads((root / "python/generativeqc/data/basis_pack.json").read_text())
    for name in ("def2-svp", "sto-3g"):
        record = {
            "name": name,
            "source": pack.get("source", "Basis Set Exchange"),
            "license": "Basis Set Exchange data",
            "elements": _pack(pack, name),
        }
        (HERE / f"{name}.json").write_text(json.dumps(record, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
