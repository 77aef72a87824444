"""Convert the reference's DE parity fixtures (RON) into JSON vectors under tests/golden/.

Source: crates/syfrack-core/tests/fixtures/parity/{sphere,mandelbulb,mandelbox,menger}.ron
of the SyFrack tree, 64 points each, written by the reference's hand formulas and checked by
crates/syfrack-core/tests/graph_parity.rs:34-82 (distance 1e-5 relative, absolute below 1;
traps 1e-5; iterations exact).  The points are LCG(1664525, 1013904223) seed 11 over the
extents sphere 3, mandelbulb 1.6, mandelbox 8, menger 1.6 (graph_parity.rs:4-17).

Only data is written: the inputs and expected outputs, with their decimal text kept
verbatim so that f32 values round-trip exactly.  Run from the repo root:
    python tests/golden/convert_de_fixtures.py /root/reference
"""
import json
import os
import re
import sys

EXTENTS = {"sphere": 3.0, "mandelbulb": 1.6, "mandelbox": 8.0, "menger": 1.6}
NUM = r"(-?[0-9.eE+-]+)"
ROW = re.compile(
    r"\(p: \(" + NUM + ", " + NUM + ", " + NUM + r"\), distance: " + NUM + r", traps: \(" + NUM + ", " + NUM
    + ", " + NUM + ", " + NUM + r"\), iterations: (\d+)\)"
)


def convert(ref_root: str, out_dir: str) -> None:
    src = os.path.join(ref_root, "crates/syfrack-core/tests/fixtures/parity")
    for name, extent in EXTENTS.items():
        rows = []
        with open(os.path.join(src, name + ".ron"), encoding="utf-8") as fh:
            for line in fh:
                m = ROW.search(line)
                if m:
                    g = m.groups()
                    rows.append({"p": list(g[0:3]), "distance": g[3], "traps": list(g[4:8]),
                                 "iterations": int(g[8])})
        assert len(rows) == 64, (name, len(rows))
        doc = {"name": name, "extent": extent, "lcg_seed": 11, "source": f"crates/syfrack-core/tests/fixtures/parity/{name}.ron",
               "tolerance": {"distance_rel": 1e-5, "traps_abs": 1e-5, "iterations": "exact"}, "samples": rows}
        with open(os.path.join(out_dir, f"de_{name}.json"), "w", encoding="utf-8") as fh:
            json.dump(doc, fh, indent=1)


if __name__ == "__main__":
    convert(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", os.path.dirname(os.path.abspath(__file__)))
