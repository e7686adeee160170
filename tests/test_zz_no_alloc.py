"""The scan entry point that allocates nothing (osh_offset_scans_noalloc)
against osh_offset_scans, through the C ABI on device pointers.

Two calls in a row, the second with smaller inputs, each compared with the
allocating entry point's output and with numpy's running sums (bit-exact
integers); sizes span empty jobs, single tiles and multi-tile look-backs.
The scratch keeps the last call's status words, so the second call also
shows that stale words never leak into a new scan.  A request beyond the
scratch's 16384 tiles is refused on the host before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def scans(ctx, arrays, noalloc):
    from omega_h_amd.hip import DeviceArray
    ins = [DeviceArray(a) for a in arrays]
    outs = [DeviceArray(dtype=np.int32, n=len(a) + 1) for a in arrays]
    f = ctx.offset_scans_noalloc if noalloc else ctx.offset_scans
    f([d.ptr for d in ins], [len(a) for a in arrays], [d.ptr for d in outs])
    ctx.synchronize()
    return [o.numpy() for o in outs]


def expected(a):
    return np.concatenate([[0], np.cumsum(a, dtype=np.int64)]).astype(np.int32)


def test_noalloc_scans_match_the_allocating_entry_point(ctx):
    rng = np.random.default_rng(3)
    for sizes in ([3_000_000, 0, 4097, 250_000, 17], [1_000_000, 5, 4096]):
        arrays = [rng.integers(-4, 9, size=n, dtype=np.int32) for n in sizes]
        print(f"noalloc scans {sizes}", flush=True)
        want = scans(ctx, arrays, False)
        got = scans(ctx, arrays, True)
        for a, w, g in zip(arrays, want, got):
            assert np.array_equal(g, w), sizes
            assert np.array_equal(g, expected(a)), sizes


def test_noalloc_scans_repeat_on_a_second_context(ctx):
    other = A.Context(0)
    rng = np.random.default_rng(5)
    for k in range(6):
        a = rng.integers(0, 3, size=int(rng.integers(1, 2_000_000)), dtype=np.int32)
        (got,) = scans(other if k % 2 else ctx, [a], True)
        assert np.array_equal(got, expected(a)), k


def test_noalloc_refuses_more_than_its_scratch(ctx):
    from omega_h_amd.hip import DeviceArray
    n = 16384 * 4096 + 1
    a = DeviceArray(dtype=np.int32, n=n)
    o = DeviceArray(dtype=np.int32, n=n + 1)
    with pytest.raises(A.OshError):
        ctx.offset_scans_noalloc([a.ptr], [n], [o.ptr])
