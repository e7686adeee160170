"""Shared helpers for the per-pass adapt() digest fixtures in tests/golden/."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

KIND = {0: "refine", 1: "coarsen", 2: "swap", 3: "sliver"}
PART = ("topology/classification", "coordinates", "metric", "lengths/qualities")


def load(name):
    with open(os.path.join(GOLDEN, f"adapt_{name}.json")) as f:
        return json.load(f)


def rows_of(passes, digests):
    return [[k, a, b] + ["%016x" % d for d in dg] for (k, a, b, _), dg in zip(passes, digests)]


def first_difference(expected, got):
    """None when the pass records agree, else a message naming the first pass
    and the part of the mesh that differs."""
    for i, (e, g) in enumerate(zip(expected, got)):
        if e[:3] != g[:3]:
            return (f"pass {i}: expected {KIND[e[0]]} {e[1]}->{e[2]} elements, "
                    f"got {KIND[g[0]]} {g[1]}->{g[2]}")
        for k in range(4):
            if e[3 + k] != g[3 + k]:
                return f"pass {i} ({KIND[e[0]]}): {PART[k]} digest differs"
    if len(expected) != len(got):
        return f"{len(expected)} passes expected, {len(got)} run"
    return None
