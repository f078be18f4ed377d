"""Writes support_matrix.json: the reference's function-support matrix.

The reference's functional test runs 32 tuning records per data type on the
problem M=768, N=640, K=256 (test/gemm_test.sh:5-73, whose -m/-n flags set the
MFMA tiles per wave along M/N in the old flag meaning: the r/s of
src/gemmMfma.cpp:44-45), and records which ones its kernel writer supports,
with the reason for each refusal, in test/result data/fuction_test.xlsx (one
sheet per data type: fp32 first, then fp16).  This script reads that
spreadsheet as data (an .xlsx is a zip of XML; no spreadsheet library is
needed) and stores the table as JSON, so the tests carry it without the
reference tree.

Run: python tests/golden/make_support_matrix.py
"""
from __future__ import annotations

import json
import os
import re
import zipfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/test/result data/fuction_test.xlsx"
OUT = os.path.join(HERE, "support_matrix.json")


def _cells(xml: str, strings: list[str]):
    for row in re.findall(r"<row[^>]*>(.*?)</row>", xml, re.S):
        vals = {}
        for col, attr, v in re.findall(r'<c r="([A-Z]+)\d+"([^>]*)>(?:<f>[^<]*</f>)?(?:<v>([^<]*)</v>)?', row):
            if 't="s"' in attr and v:
                v = strings[int(v)]
            vals[col] = v
        yield vals


def main() -> None:
    with zipfile.ZipFile(SRC) as z:
        strings = re.findall(r"<t[^>]*>([^<]*)</t>", z.read("xl/sharedStrings.xml").decode())
        sheets = {}
        for name, dtype in (("sheet1", 1), ("sheet2", 2)):
            rows = []
            problem = None
            for c in _cells(z.read(f"xl/worksheets/{name}.xml").decode(), strings):
                if c.get("B") == "M":
                    problem = {"M": int(c["C"]), "N": int(c["E"]), "K": int(c["G"])}
                if c.get("B", "").isdigit() and c.get("H") in ("Y", "N"):
                    rows.append({"index": int(c["B"]), "r": int(c["C"]), "s": int(c["D"]), "x": int(c["E"]),
                                 "y": int(c["F"]), "u": int(c["G"]), "support": c["H"],
                                 "reason": c.get("I", "").strip()})
            sheets[str(dtype)] = {"problem": problem, "configs": rows}
    doc = {"source": "test/result data/fuction_test.xlsx (sheet1: -d 1 fp32, sheet2: -d 2 fp16); "
                     "the records are test/gemm_test.sh:5-73",
           "columns": "r, s: MFMA tiles per wave along M, N; x, y: waves along M, N; u: DepthU",
           "dtypes": sheets}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(f"wrote {OUT}: " + ", ".join(f"dtype {k}: {len(v['configs'])} configs" for k, v in sheets.items()))


if __name__ == "__main__":
    main()
