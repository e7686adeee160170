"""Named ranges (Omega_h::profile's timer tree, src/Omega_h_profile.hpp:47-232).

On the CPU: the tree and its TOP-DOWN/BOTTOM-UP listings on a fixed sequence
of starts, stops and clock readings, against listings worked out by hand
(tests/cpp/profile_tree.cpp).  On the GPU: ranges around a real adapt() form
the reference's call tree, cover every pass and leave the mesh unchanged."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "omega_h_amd", "csrc")


def test_timer_tree_listings(tmp_path):
    exe = str(tmp_path / "profile_tree")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + CSRC,
                    os.path.join(ROOT, "tests", "cpp", "profile_tree.cpp"),
                    os.path.join(CSRC, "osh_profile.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def children(frames, parent):
    return {f[0]: i for i, f in enumerate(frames) if f[1] == parent}


@pytest.mark.gpu
def test_adapt_ranges_form_the_call_tree():
    A = pytest.importorskip("omega_h_amd")
    from omega_h_amd import workloads as W
    ctx = A.Context(0)
    opts = ctx.default_opts(3)
    opts.verbosity = 0

    def make():
        m = A.Mesh.box(ctx, 3, 6)
        met, nc = W.levelset_metric_3d(m.coords(), h_min=0.02, h_max=0.25, width=0.1)
        m.set_metric(met, nc)
        return m

    plain = make()
    plain.adapt(opts)
    ctx.set_profile(A.Context.PROFILE_TREE | A.Context.PROFILE_SYNC)
    m = make()
    did, passes = m.adapt(opts)
    ctx.set_profile(0)
    assert did and m.digests() == plain.digests()
    fr = ctx.profile_frames()
    roots = children(fr, -1)
    assert "adapt" in roots
    top = children(fr, roots["adapt"])
    assert {"pre_adapt", "satisfy_lengths", "satisfy_quality"} <= set(top)
    kinds = np.bincount([p[0] for p in passes], minlength=4)
    under = {**children(fr, top["satisfy_lengths"]), **children(fr, top["satisfy_quality"])}
    names = ["refine_by_size", "coarsen_by_size", "swap_edges", "coarsen_slivers"]
    for k, name in enumerate(names):
        if kinds[k]:
            # every pass that changed the mesh opened its range (so did the
            # attempts that changed nothing)
            calls = sum(fr[i][3] for i, f in enumerate(fr) if f[0] == name)
            assert calls >= kinds[k], name
    assert "coarsen_by_size" in under
    # stages of a pass nest under it
    cbs = under["coarsen_by_size"]
    assert {"find_indset", "modify_ents"} <= set(children(fr, cbs))
    # a parent's time covers its children's
    for i, (name, parent, secs, calls) in enumerate(fr):
        kids = [f[2] for f in fr if f[1] == i]
        assert secs >= 0 and calls >= 1
        assert sum(kids) <= secs * (1 + 1e-9) + 1e-6, name
    rep = ctx.profile_report(chop=1.0, percent=True)
    assert "TOP-DOWN:" in rep and "BOTTOM-UP:" in rep and "adapt" in rep
    # markers only: no tree kept, same mesh
    ctx.profile_clear()
    ctx.set_profile(A.Context.PROFILE_MARKERS)
    m2 = make()
    m2.adapt(opts)
    ctx.set_profile(0)
    assert ctx.profile_frames() == [] and m2.digests() == plain.digests()
