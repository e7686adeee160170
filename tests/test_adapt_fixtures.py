"""The per-pass oracle digests in tests/golden/ (written by
tests/golden/make_adapt_digests.py) still describe what the oracle does: the
small workloads are run again and must give the same record, pass by pass.
The 64^3 file (two minutes of oracle time) is checked for its shape only."""
import importlib.util
import os

import pytest

import adapt_fixtures as F

spec = importlib.util.spec_from_file_location(
    "make_adapt_digests", os.path.join(F.GOLDEN, "make_adapt_digests.py"))
G = importlib.util.module_from_spec(spec)
spec.loader.exec_module(G)


@pytest.mark.parametrize("name", ["ring_2d_32", "levelset_3d_8"])
def test_oracle_reproduces_fixture(name):
    fx = F.load(name)
    rec = G.run(name)
    assert rec["final_nents"] == fx["final_nents"]
    msg = F.first_difference(fx["passes"], rec["passes"])
    assert msg is None, msg


@pytest.mark.parametrize("name", list(G.WORKLOADS))
def test_fixture_shape(name):
    fx = F.load(name)
    rows = fx["passes"]
    assert fx["adapted"] == 1 and len(rows) > 0
    # element counts chain from pass to pass
    for a, b in zip(rows, rows[1:]):
        assert a[2] == b[1]
    assert rows[-1][2] == fx["final_nents"][-1]
    d = fx["dim"]
    n = fx["cells_per_side"]
    assert rows[0][1] == (6 * n ** 3 if d == 3 else 2 * n * n)
    for r in rows:
        assert r[0] in (0, 1, 2, 3)
        assert all(len(h) == 16 for h in r[3:])
