"""The C++ drop-in facade (include/Omega_h_amd.hpp).

CPU: the header compiles on its own.  GPU: C++ programs written the way an
Omega_h caller writes adapt() (tests/cpp/facade_*.cpp, built by
__graft_entry__.build()) give bit-for-bit the oracle's adapted mesh, and the
facade's arrays live on the device (tests/cpp/facade_device.cpp).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "omega_h_amd", "_build", "facade_adapt")


@pytest.mark.parametrize("std", ["c++14", "c++17"])
def test_facade_header_compiles(tmp_path, std):
    src = tmp_path / "use.cpp"
    src.write_text('#include "Omega_h_amd.hpp"\n'
                   "int main() { Omega_h::AdaptOpts o; (void)o; return 0; }\n")
    r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("every", [False, True], ids=["defaults", "every_field"])
@pytest.mark.parametrize("dim,n", [(2, 16), (3, 6)])
def test_facade_adapt_matches_oracle(tmp_path, dim, n, every):
    from omega_h_amd import workloads as W
    assert os.path.exists(BIN), "build the facade driver with __graft_entry__.build()"
    o = O.Mesh.box(dim, n)
    met, nc = W.iso_ring_metric_2d(o.coords()) if dim == 2 else W.levelset_metric_3d(o.coords())
    mfile = tmp_path / "metric.f64"
    np.ascontiguousarray(met, dtype=np.float64).tofile(mfile)
    prefix = str(tmp_path / "out")
    print(f"facade adapt dim={dim} n={n} every={every}", flush=True)
    r = subprocess.run([BIN, str(dim), str(n), str(nc), str(mfile), prefix] +
                       (["every"] if every else []),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o.set_metric(met, nc)
    oo = O.default_opts(dim)
    oo.should_prevent_coarsen_flip = int(every)
    did, _ = o.adapt(oo)
    counts = np.fromfile(prefix + ".counts", dtype=np.int32)
    assert list(counts[:-1]) == [o.nents(d) for d in range(dim + 1)]
    assert counts[-1] == did
    assert np.array_equal(np.fromfile(prefix + ".verts", dtype=np.int32), o.verts(dim))
    for name, want in (("coords", o.coords()), ("quality", o.qualities())):
        got = np.fromfile(prefix + "." + name, dtype=np.float64)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name


@pytest.mark.gpu
def test_facade_approach_loop_matches_oracle(tmp_path):
    prog = os.path.join(ROOT, "omega_h_amd", "_build", "facade_approach")
    assert os.path.exists(prog), "build the facade programs with __graft_entry__.build()"
    o = O.Mesh.box(2, 8)
    c = o.coords().reshape(-1, 2)
    init = np.full(o.nents(0), 64.0)
    target = 1.0 / (0.03 + 0.1 * c[:, 0]) ** 2
    init.tofile(tmp_path / "m.f64")
    target.tofile(tmp_path / "t.f64")
    prefix = str(tmp_path / "out")
    print("facade approach 2D n=8", flush=True)
    r = subprocess.run([prog, "2", "8", "1", str(tmp_path / "m.f64"), str(tmp_path / "t.f64"),
                        prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o.set_metric(init, 1)
    o.add_vertex_tag("target_metric", 1, O.Mesh.METRIC, target)
    steps = 0
    while o.approach_metric(O.default_opts(2)) > 0:
        o.adapt(O.default_opts(2))
        steps += 1
    counts = np.fromfile(prefix + ".counts", dtype=np.int32)
    assert list(counts[:-1]) == [o.nents(d) for d in range(3)]
    assert counts[-1] == steps
    assert np.array_equal(np.fromfile(prefix + ".verts", dtype=np.int32), o.verts(2))
    for name, want in (("coords", o.coords()), ("metric", o.metric())):
        got = np.fromfile(prefix + "." + name, dtype=np.float64)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name


@pytest.mark.gpu
def test_facade_warp_loop_matches_oracle(tmp_path):
    """do adapt while warp_to_limit (src/bend_test.cpp:55-65) through the
    facade, then ask_sizes and ask_dual, against the oracle bit for bit."""
    prog = os.path.join(ROOT, "omega_h_amd", "_build", "facade_warp")
    assert os.path.exists(prog), "build the facade programs with __graft_entry__.build()"
    o = O.Mesh.box(3, 3)
    c = o.coords().reshape(-1, 3)
    met = np.full(o.nents(0), 9.0)
    w = np.zeros_like(c)
    w[:, 0] = 2.0 * c[:, 2] ** 2
    met.tofile(tmp_path / "m.f64")
    w.tofile(tmp_path / "w.f64")
    prefix = str(tmp_path / "out")
    print("facade warp 3D n=3", flush=True)
    r = subprocess.run([prog, "3", "3", "1", str(tmp_path / "m.f64"), str(tmp_path / "w.f64"),
                        prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o.set_metric(met, 1)
    o.add_vertex_tag("warp", 3, O.Mesh.LINEAR, w.ravel())
    steps = 0
    while True:
        o.adapt(O.default_opts(3))
        steps += 1
        if o.warp_to_limit(O.default_opts(3)) != 1:
            break
    counts = np.fromfile(prefix + ".counts", dtype=np.int32)
    assert list(counts[:-1]) == [o.nents(d) for d in range(4)]
    assert counts[-1] == steps
    assert np.array_equal(np.fromfile(prefix + ".verts", dtype=np.int32), o.verts(3))
    for name, want in (("coords", o.coords()), ("sizes", o.sizes())):
        got = np.fromfile(prefix + "." + name, dtype=np.float64)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name
    a2ab, ab2b = o.dual()
    assert np.array_equal(np.fromfile(prefix + ".dual", dtype=np.int32),
                          np.concatenate([a2ab, ab2b]))


def sequential_centroids(coords, ev2v, dim):
    """Element centroids summed vertex by vertex from 0.0, then divided, as
    average_field (src/Omega_h_mesh.cpp:822-844) forms them."""
    c = coords.reshape(-1, dim)
    ev = ev2v.reshape(-1, dim + 1)
    s = np.zeros((len(ev), dim))
    for k in range(dim + 1):
        s = s + c[ev[:, k]]
    return s / (dim + 1)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", [(2, 10), (3, 6)])
def test_facade_conserve_matches_oracle(tmp_path, dim, n):
    """src/2d_conserve_test.cpp and src/3d_conserve_test.cpp through the
    facade's TransferOpts maps, against the oracle bit for bit."""
    prog = os.path.join(ROOT, "omega_h_amd", "_build", "facade_conserve")
    assert os.path.exists(prog), "build the facade programs with __graft_entry__.build()"
    o = O.Mesh.box(dim, n)
    metric = o.implied_isos() * (1.0 / (1.3 * 1.3))
    metric.tofile(tmp_path / "m.f64")
    prefix = str(tmp_path / "out")
    print(f"facade conserve dim={dim} n={n}", flush=True)
    r = subprocess.run([prog, str(dim), str(n), str(tmp_path / "m.f64"), prefix],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o.set_metric(metric, 1)
    cen = sequential_centroids(o.coords(), o.verts(dim), dim)
    o.add_elem_tag("density", 1, np.where(cen[:, 1] > 0.5, 1e-4, 1.0))
    c = o.coords().reshape(-1, dim)
    x2 = c[:, 0] - 0.3
    vel = np.zeros((len(c), dim))
    vel[:, 0] = np.sign(x2) * np.sqrt(np.abs(x2))
    o.add_vertex_tag("velocity", dim, O.Mesh.LINEAR, vel.ravel())
    o.set_elem_transfer("density", O.Mesh.CONSERVE, (1, 0.9, 0.0))
    o.set_momentum_velocity("velocity", "density", (1, 0.02, 1e-6))
    opts = O.default_opts(dim)
    opts.verbosity = 0
    o.adapt(opts)
    counts = np.fromfile(prefix + ".counts", dtype=np.int32)
    assert list(counts) == [o.nents(d) for d in range(dim + 1)]
    assert np.array_equal(np.fromfile(prefix + ".verts", dtype=np.int32), o.verts(dim))
    for name, want in (("coords", o.coords()), ("density", o.elem_tag("density")),
                       ("velocity", o.vertex_tag("velocity"))):
        got = np.fromfile(prefix + "." + name, dtype=np.float64)
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), name


@pytest.mark.gpu
def test_facade_arrays_live_on_the_device():
    """Read/Write of the facade are device memory: tags round-trip through
    device-to-device copies, fills, ranges, single items and shared storage
    behave as the reference's arrays, device reductions equal host ones, and
    an array outlives its Library (tests/cpp/facade_device.cpp)."""
    prog = os.path.join(ROOT, "omega_h_amd", "_build", "facade_device")
    assert os.path.exists(prog), "build the facade programs with __graft_entry__.build()"
    print("facade device arrays", flush=True)
    r = subprocess.run([prog], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok"), r.stdout
