"""The .osh binary format (src/Omega_h_file.cpp:88-600): round trips through
the GPU mesh, and a reader restated here from the reference's layout
(magic, compression flag, meta, nverts, down adjacencies, tags per
dimension, class sets, parents) that decodes what the library writes and
writes what the library reads."""
import os
import struct
import zlib

import numpy as np
import pytest

A = pytest.importorskip("omega_h_amd")
from omega_h_amd import workloads as W  # noqa: E402

TYPES = {0: np.int8, 2: np.int32, 3: np.int64, 5: np.float64}


class Reader:
    def __init__(self, data):
        self.d = data
        self.i = 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.d, self.i)
        self.i += struct.calcsize("<" + fmt)
        return v[0] if len(v) == 1 else v

    def array(self, dtype, compressed):
        n = self.take("i")
        nbytes = n * np.dtype(dtype).itemsize
        if compressed:
            zb = self.take("q")
            raw = zlib.decompress(self.d[self.i:self.i + zb])
            self.i += zb
        else:
            raw = self.d[self.i:self.i + nbytes]
            self.i += nbytes
        assert len(raw) == nbytes
        return np.frombuffer(raw, dtype=dtype)


def decode(path):
    assert open(os.path.join(path, "nparts")).read().strip() == "1"
    version = int(open(os.path.join(path, "version")).read())
    r = Reader(open(os.path.join(path, "0.osh"), "rb").read())
    assert r.take("BB") == (0xA1, 0x1A)
    z = r.take("b") != 0
    family, dim, size, rank, parting, nghost, hints = r.take("bbiibib")
    assert (family, size, rank, hints) == (0, 1, 0, 0)
    out = {"version": version, "dim": dim, "nverts": r.take("i"), "down": {}, "codes": {},
           "tags": {}}
    for d in range(1, dim + 1):
        out["down"][d] = r.array(np.int32, z)
        if d > 1:
            out["codes"][d] = r.array(np.int8, z)
    for d in range(dim + 1):
        for _ in range(r.take("i")):
            name = r.d[r.i + 4:r.i + 4 + r.take("i")].decode()
            r.i += len(name)
            ncomps, typ = r.take("bb")
            out["tags"][(d, name)] = (ncomps, r.array(TYPES[typ], z))
    assert r.take("i") == 0 and r.take("b") == 0
    assert r.i == len(r.d)
    return out


def encode(path, dim, nverts, down, codes, tags):
    """An uncompressed .osh directory in the reference's layout."""
    os.makedirs(path, exist_ok=True)
    open(os.path.join(path, "nparts"), "w").write("1\n")
    open(os.path.join(path, "version"), "w").write("9\n")
    b = bytearray(b"\xa1\x1a") + struct.pack("<b", 0)
    b += struct.pack("<bbiibib", 0, dim, 1, 0, 0, 0, 0) + struct.pack("<i", nverts)

    def arr(a):
        return struct.pack("<i", len(a)) + np.ascontiguousarray(a).tobytes()

    for d in range(1, dim + 1):
        b += arr(down[d])
        if d > 1:
            b += arr(codes[d])
    for d in range(dim + 1):
        mine = [(n, v) for (dd, n), v in tags.items() if dd == d]
        b += struct.pack("<i", len(mine))
        for name, (ncomps, a) in mine:
            typ = {np.dtype(t): k for k, t in TYPES.items()}[a.dtype]
            b += struct.pack("<i", len(name)) + name.encode() + struct.pack("<bb", ncomps, typ)
            b += arr(a)
    b += struct.pack("<ib", 0, 0)
    open(os.path.join(path, "0.osh"), "wb").write(bytes(b))


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n", [(2, 10), (3, 4)])
def test_write_matches_the_reference_layout_and_round_trips(tmp_path, dim, n):
    ctx = A.Context(0)
    g = A.Mesh.box(ctx, dim, n)
    met, nc = W.iso_ring_metric_2d(g.coords()) if dim == 2 else W.levelset_metric_3d(g.coords())
    g.set_metric(met, nc)
    g.add_vertex_tag("xyz", dim, 0, g.coords())
    g.adapt(ctx.default_opts(dim))
    path = str(tmp_path / "mesh.osh")
    g.write_osh(path)
    f = decode(path)
    assert f["version"] == 9 and f["dim"] == dim and f["nverts"] == g.nents(0)
    for d in range(1, dim + 1):
        down, codes = g.down(d)
        assert np.array_equal(f["down"][d], down)
        if d > 1:
            assert np.array_equal(f["codes"][d], codes)
    assert np.array_equal(f["tags"][(0, "coordinates")][1], g.coords())
    assert np.array_equal(f["tags"][(0, "metric")][1], g.metric())
    assert np.array_equal(f["tags"][(0, "xyz")][1], g.coords())
    for d in range(dim + 1):
        ids, dims = g.classification(d)
        assert np.array_equal(f["tags"][(d, "class_id")][1], ids)
        assert np.array_equal(f["tags"][(d, "class_dim")][1], dims)
        assert np.array_equal(f["tags"][(d, "global")][1], np.arange(g.nents(d)))
    # read back: the same mesh, which then adapts identically
    h = A.Mesh.read_osh(ctx, path)
    assert h.digest() == g.digest()
    assert np.array_equal(h.metric(), g.metric())
    assert np.array_equal(h.vertex_tag("xyz"), g.vertex_tag("xyz"))
    opts = ctx.default_opts(dim)
    opts.min_quality_desired = 0.45 if dim == 2 else 0.35
    g.adapt(opts)
    h.adapt(opts)
    assert h.digest() == g.digest()


@pytest.mark.gpu
def test_read_an_uncompressed_file_written_in_the_reference_layout(tmp_path):
    ctx = A.Context(0)
    g = A.Mesh.box(ctx, 3, 2)
    tags = {(0, "coordinates"): (3, g.coords())}
    for d in range(4):
        ids, dims = g.classification(d)
        tags[(d, "class_id")] = (1, ids)
        tags[(d, "class_dim")] = (1, dims)
        tags[(d, "global")] = (1, np.arange(g.nents(d), dtype=np.int64))
    tags[(0, "target_metric")] = (1, np.full(g.nents(0), 4.0))
    down = {d: g.down(d)[0] for d in (1, 2, 3)}
    codes = {d: g.down(d)[1] for d in (2, 3)}
    path = str(tmp_path / "ref.osh")
    encode(path, 3, g.nents(0), down, codes, tags)
    h = A.Mesh.read_osh(ctx, path)
    assert h.digest() == g.digest()
    assert np.array_equal(h.vertex_tag("target_metric"), np.full(g.nents(0), 4.0))
    for d in range(4):
        for x, y in zip(h.classification(d), g.classification(d)):
            assert np.array_equal(x, y)


def test_read_rejects_what_is_not_an_osh_directory(tmp_path):
    try:
        ctx = A.Context(0)
    except A.OshError:
        pytest.skip("no GPU")
    os.makedirs(tmp_path / "bad.osh")
    (tmp_path / "bad.osh" / "nparts").write_text("1\n")
    (tmp_path / "bad.osh" / "version").write_text("9\n")
    (tmp_path / "bad.osh" / "0.osh").write_bytes(b"\x00\x00\x00")
    with pytest.raises(A.OshError, match="not an .osh file"):
        A.Mesh.read_osh(ctx, str(tmp_path / "bad.osh"))
