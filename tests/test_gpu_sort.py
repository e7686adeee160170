"""sort_by_keys on the device (the own stable LSD radix sort of
osh_prims.hip): the reference's known answers (src/unit_array_algs.cpp:
36-57) and, for what they leave open, the permutation numpy's stable
lexical sort gives (np.lexsort sorts by the last key first and is stable, so
it is the stable tuple order the reference's sort_by_keys defines) on random
tuples with many ties, negative keys, keys that differ only in their high
bytes, and ragged sizes around the 4096-key tile."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


@pytest.mark.parametrize("keys,width,want", [
    ([0, 1], 1, [0, 1]),
    ([0, 2, 0, 1], 2, [1, 0]),
    ([0, 2, 1, 1], 2, [0, 1]),
    ([1, 2, 3, 1, 2, 2, 3, 0, 0], 3, [1, 0, 2]),
])
def test_reference_known_answers(ctx, keys, width, want):
    assert ctx.sort_by_keys(keys, width).tolist() == want


def stable_order(keys, width):
    k = np.asarray(keys, dtype=np.int64).reshape(-1, width)
    # lexsort takes the primary key last
    return np.lexsort([k[:, j] for j in range(width - 1, -1, -1)]).astype(np.int32)


@pytest.mark.parametrize("n", [0, 1, 2, 4095, 4096, 4097, 100003])
@pytest.mark.parametrize("width", [1, 2, 3])
def test_random_tuples_match_a_stable_lexical_sort(ctx, n, width):
    rng = np.random.default_rng(1000 * n + width)
    # few distinct values per column, so that ties are common and stability
    # decides the order
    k = rng.integers(-3, 4, size=(n, width)).astype(np.int64)
    print(f"sort n={n} width={width}", flush=True)
    assert np.array_equal(ctx.sort_by_keys(k.ravel(), width), stable_order(k.ravel(), width))


def test_keys_differing_in_high_bytes_only(ctx):
    """Digits that are equal across all keys are skipped; keys that differ
    only above bit 32, or only in the sign, still sort (signed order)."""
    rng = np.random.default_rng(7)
    n = 50000
    k = (rng.integers(-2**20, 2**20, size=n).astype(np.int64) << 40)
    k[::7] = np.iinfo(np.int64).min
    k[::11] = np.iinfo(np.int64).max
    assert np.array_equal(ctx.sort_by_keys(k, 1), stable_order(k, 1))
    same = np.full(n, 12345, dtype=np.int64)
    assert np.array_equal(ctx.sort_by_keys(same, 1), np.arange(n, dtype=np.int32))
