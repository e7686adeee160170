"""The batched single-pass offset scan (offset_scan, src/Omega_h_int_scan.cpp:10-20)
against numpy's running sums, through the C ABI on device pointers.

Sizes cover empty arrays, one partial tile, the 4096-value tile boundary and
scans of hundreds to thousands of tiles, alone and mixed in one launch, so the
look-back crosses windows of 64 tiles and each job's look-back stays inside
its own tiles.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def run_scans(ctx, arrays):
    from omega_h_amd.hip import DeviceArray
    ins = [DeviceArray(a) for a in arrays]
    outs = [DeviceArray(dtype=np.int32, n=len(a) + 1) for a in arrays]
    ctx.offset_scans([d.ptr for d in ins], [len(a) for a in arrays], [d.ptr for d in outs])
    ctx.synchronize()
    return [o.numpy() for o in outs]


def expected(a):
    return np.concatenate([[0], np.cumsum(a, dtype=np.int64)]).astype(np.int32)


@pytest.mark.parametrize("n", [0, 1, 17, 4095, 4096, 4097, 8192, 100_003, 262_144 + 5,
                               3_000_000, 12_000_000])
def test_single_scan(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 8, size=n, dtype=np.int32)
    (got,) = run_scans(ctx, [a])
    assert np.array_equal(got, expected(a)), n


def test_batched_scans_in_one_launch(ctx):
    rng = np.random.default_rng(7)
    for sizes in ([5, 0, 4096, 1], [300_000, 17, 1_000_000, 4097], [2_500_000, 2_500_000],
                  [0, 0, 0], [1000, 5000, 0, 70_000, 4096, 3, 250_000, 9000],
                  [7, 300_000, 0, 4096, 50_000, 1, 200_000, 4097, 12, 100_000, 0, 65_536]):
        arrays = [rng.integers(0, 5, size=n, dtype=np.int32) for n in sizes]
        got = run_scans(ctx, arrays)
        for a, g in zip(arrays, got):
            assert np.array_equal(g, expected(a)), sizes


def test_repeated_scans_keep_the_status_words_apart(ctx):
    # consecutive launches reuse the status words under new epochs
    rng = np.random.default_rng(11)
    for k in range(20):
        a = rng.integers(-3, 4, size=int(rng.integers(1, 600_000)), dtype=np.int32)
        (got,) = run_scans(ctx, [a])
        assert np.array_equal(got, expected(a)), k


def test_bad_scan_arguments_are_refused(ctx):
    from omega_h_amd.hip import DeviceArray
    a = DeviceArray(np.ones(4, np.int32))
    o = DeviceArray(dtype=np.int32, n=5)
    with pytest.raises(A.OshError):
        ctx.offset_scans([a.ptr] * 13, [4] * 13, [o.ptr] * 13)
    with pytest.raises(A.OshError):
        ctx.offset_scans([a.ptr], [-1], [o.ptr])
