"""Write tests/golden/qrs45_n1to4.json: the Spence QRS45 product-quadrature
tables for orders N = 1..4 as the reference tabulates them
(src/teton/mods/QuadratureData_mod.F90:1144-2857: `first`, `last`,
cosTheta, weightTheta, cosPhiXYZ, weightPhiXYZ, and the RZ azimuthal tables
cosPhiRZ, weightPhiRZ that rt/quadrz.F90 uses). Reads the reference source
as text; run once where /root/reference exists:

    python tests/golden/make_qrs45_fixture.py /root/reference
"""
import json
import os
import re
import sys


def data_block(text, name):
    m = re.search(r"data\s+" + name + r"\s*/(.*?)/", text, re.S | re.I)
    if not m:
        raise SystemExit(f"data statement {name} not found")
    body = re.sub(r"!.*", "", m.group(1))
    vals = re.findall(r"[-+]?\d+\.?\d*(?:[eEdD][-+]?\d+)?(?=_adqt|\s*[,&/]|\s*$)", body)
    return [float(v.replace("d", "e").replace("D", "e")) for v in vals]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(ref, "src/teton/mods/QuadratureData_mod.F90")
    text = open(path).read()
    first = [int(x) for x in data_block(text, "first")]
    last = [int(x) for x in data_block(text, "last")]
    tables = {k: data_block(text, k) for k in ("cosTheta", "weightTheta", "cosPhiXYZ",
                                               "weightPhiXYZ", "cosPhiRZ", "weightPhiRZ")}
    out = {"source": "src/teton/mods/QuadratureData_mod.F90 (data first/last/cosTheta/"
                     "weightTheta/cosPhiXYZ/weightPhiXYZ/cosPhiRZ/weightPhiRZ)", "orders": {}}
    for n in range(1, 5):
        lo, hi = first[n - 1], last[n - 1]
        out["orders"][str(n)] = {k: v[lo - 1:hi] for k, v in tables.items()}
        assert all(len(v) == n for v in out["orders"][str(n)].values()), n
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qrs45_n1to4.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
