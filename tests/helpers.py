"""Shared test helpers: fixture loading and the reference's LCG point corpus."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lcg_points(n: int, extent: float, seed: int) -> np.ndarray:
    """graph_parity.rs:4-13: LCG(1664525, 1013904223), (s >> 8) / 2^24, scaled to
    [-extent, extent) as lcg * (2 * extent) - extent, all in f32."""
    s = seed
    out = np.zeros((n, 3), np.float32)
    two_e = np.float32(2.0) * np.float32(extent)
    for i in range(n):
        for k in range(3):
            s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
            v = np.float32(s >> 8) / np.float32(16777216.0)
            out[i, k] = np.float32(v * two_e) - np.float32(extent)
    return out


def load_de_fixture(name: str):
    with open(os.path.join(GOLDEN, f"de_{name}.json"), encoding="utf-8") as fh:
        doc = json.load(fh)
    s = doc["samples"]
    p = np.array([[np.float32(float(v)) for v in r["p"]] for r in s], np.float32)
    d = np.array([np.float32(float(r["distance"])) for r in s], np.float32)
    t = np.array([[np.float32(float(v)) for v in r["traps"]] for r in s], np.float32)
    it = np.array([r["iterations"] for r in s], np.uint32)
    return doc, p, d, t, it
