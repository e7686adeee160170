"""Whole adapt() runs on the GPU compared with the oracle pass by pass.

Each workload's fixture (tests/golden/adapt_<name>.json, written by the
oracle through tests/golden/make_adapt_digests.py) records every
mesh-changing pass: its kind, the element counts in and out, and FNV-1a
digests of the vertex lists and classification, the coordinate bits, the
metric bits and the edge length and element quality bits.  The device's
adapt() (osh_adapt_digests, the same loop as osh_adapt with the swap choices
carried between swap passes) must reproduce every record bit for bit, in
FP64 and in the mixed-precision form (FP32 cavity screens with every cavity
not rejected by the screening band re-evaluated in FP64).
"""
import pytest

import adapt_fixtures as F

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402

METRICS = {"ring_2d_32": (2, 32, W.iso_ring_metric_2d),
           "levelset_3d_8": (3, 8, W.levelset_metric_3d),
           "levelset_3d_12": (3, 12, W.levelset_metric_3d)}


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", list(METRICS))
def test_adapt_matches_oracle_pass_by_pass(ctx, name, precision):
    dim, n, metric = METRICS[name]
    fx = F.load(name)
    m = A.Mesh.box(ctx, dim, n)
    met, nc = metric(m.coords())
    m.set_metric(met, nc)
    opts = ctx.default_opts(dim)
    opts.verbosity = 0
    opts.precision = precision
    did, passes, digests = m.adapt_digests(opts)
    assert did == fx["adapted"]
    msg = F.first_difference(fx["passes"], F.rows_of(passes, digests))
    assert msg is None, f"{name}: {msg}"
    assert [m.nents(d) for d in range(dim + 1)] == fx["final_nents"]
