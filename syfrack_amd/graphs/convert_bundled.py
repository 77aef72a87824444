"""Converts the forms shipped with the reference (graphs/*.ron of the SyFrack tree, embedded by
crates/syfrack-core/src/scene/graph/bundled.rs) into the JSON that Graph.bundled reads.

Only the graph data is kept: node ids, kinds with their parameters, names, enabled flags,
links, safety and bounding-radius hint (the editor positions are dropped).  Run from the repo
root:
    python syfrack_amd/graphs/convert_bundled.py /root/reference
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from syfrack_amd.graph import Graph  # noqa: E402


def convert(ref_root: str) -> None:
    src = os.path.join(ref_root, "graphs")
    for f in sorted(os.listdir(src)):
        if not f.endswith(".ron"):
            continue
        with open(os.path.join(src, f), encoding="utf-8") as fh:
            g = Graph.from_ron(fh.read())
        doc = {"name": f[:-4], "source": f"graphs/{f}", **g.to_dict()}
        with open(os.path.join(HERE, f[:-4] + ".json"), "w", encoding="utf-8") as fh:
            fh.write(dumps(doc) + "\n")


def dumps(doc: dict) -> str:
    """JSON with one node per line and the links on one line."""
    one = lambda v: json.dumps(v, ensure_ascii=False)  # noqa: E731
    head = ",\n".join(f" {one(k)}: {one(v)}" for k, v in doc.items() if k not in ("nodes", "links"))
    nodes = ",\n".join("  " + one(n) for n in doc["nodes"])
    return "{\n" + head + ',\n "nodes": [\n' + nodes + '\n ],\n "links": ' + one(doc["links"]) + "\n}"


if __name__ == "__main__":
    convert(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
