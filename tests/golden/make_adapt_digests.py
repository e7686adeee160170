"""Write the per-pass oracle digests of whole adapt() runs (test fixtures).

For each workload below, the oracle (oracle/liboracle.so, the CPU restatement
of Omega_h's adapt()) runs adapt() with default AdaptOpts and records, after
every mesh-changing pass, the pass kind, the element counts in and out and
four FNV-1a digests of the mesh (vertex lists and classification; coordinate
bits; metric bits; edge length and element quality bits).  The GPU tests
compare the device's run against these files pass by pass.

    python tests/golden/make_adapt_digests.py            # every workload
    python tests/golden/make_adapt_digests.py ring_2d_32 # one of them

The 64^3 level-set workload takes about two minutes of one core.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from omega_h_amd import workloads as W  # noqa: E402

WORKLOADS = {
    # 2D 32x32 box (2,048 triangles), isotropic ring h = 0.02 + 0.08|r - 0.25|
    "ring_2d_32": (2, 32, W.iso_ring_metric_2d),
    # the level-set metric (normal h 0.002..0.05 around z = 0.5 + 0.1 sin 2 pi x)
    "levelset_3d_8": (3, 8, W.levelset_metric_3d),
    "levelset_3d_12": (3, 12, W.levelset_metric_3d),
    "levelset_3d_64": (3, 64, W.levelset_metric_3d),
}


def path_of(name):
    return os.path.join(HERE, f"adapt_{name}.json")


def run(name):
    dim, n, metric = WORKLOADS[name]
    m = O.Mesh.box(dim, n)
    met, nc = metric(m.coords())
    m.set_metric(met, nc)
    opts = O.default_opts(dim)
    opts.verbosity = 0
    t0 = time.time()
    did, passes, digests = m.adapt_digests(opts)
    secs = time.time() - t0
    rows = [[k, a, b] + ["%016x" % d for d in dg] for (k, a, b, _), dg in zip(passes, digests)]
    return {
        "workload": name,
        "dim": dim,
        "cells_per_side": n,
        "adapted": int(did),
        "passes": rows,
        "final_nents": [m.nents(d) for d in range(dim + 1)],
        "oracle_seconds": round(secs, 1),
    }


def main(names):
    for name in names:
        rec = run(name)
        rows = rec.pop("passes")
        text = json.dumps(rec)[:-1] + ', "passes": [\n'
        text += ",\n".join(json.dumps(r) for r in rows) + "\n]}\n"
        with open(path_of(name), "w") as f:
            f.write(text)
        print(name, len(rows), "passes", rec["final_nents"], rec["oracle_seconds"], "s")


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
