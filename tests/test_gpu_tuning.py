"""The context's lane counts (osh_ctx_set_tuning) change how work is spread
over the device and nothing else: whole adapts of a 3D level-set box and a
2D ring, in FP64 and in the mixed form, give the same meshes (all four
digests) at the smallest and largest lane count of every knob as with the
defaults.  Invalid settings are refused."""
import pytest

pytestmark = pytest.mark.gpu

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    return A.Context(0)


def adapted(ctx, dim, n, precision):
    m = A.Mesh.box(ctx, dim, n)
    if dim == 3:
        met, nc = W.levelset_metric_3d(m.coords())
    else:
        met, nc = W.iso_ring_metric_2d(m.coords())
    m.set_metric(met, nc)
    o = ctx.default_opts(dim)
    o.verbosity = 0
    o.precision = precision
    _, passes = m.adapt(o)
    return m.digests(), [p[:3] for p in passes]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("dim,n", [(3, 12), (2, 32)])
def test_lane_counts_change_no_result(ctx, dim, n, precision):
    defaults = ctx.tuning()
    want = adapted(ctx, dim, n, precision)
    try:
        for name in ctx.TUNING:
            for v in (1, 64):
                ctx.set_tuning(name, v)
                print(f"dim {dim} n {n} precision {precision}: {name} = {v}", flush=True)
                assert adapted(ctx, dim, n, precision) == want, (name, v)
                ctx.set_tuning(name, defaults[name])
        ctx.apply_tuning({k: 2 for k in ctx.TUNING})
        assert adapted(ctx, dim, n, precision) == want
    finally:
        ctx.apply_tuning(defaults)


def test_bad_settings_are_refused(ctx):
    with pytest.raises(A.OshError):
        ctx.set_tuning("coarsen_lanes", 12)
    with pytest.raises(A.OshError):
        ctx.set_tuning("coarsen_lanes", 128)
    with pytest.raises(A.OshError):
        ctx.set_tuning("no_such_knob", 8)
