"""Synthetic adapt() workloads: box meshes and their target metric fields.

Each metric is evaluated at vertex coordinates in float64 and stored in
Omega_h's symmetric layout (src/Omega_h_matrix.hpp:391-431):
1 component (isotropic 1/h^2), 3 components (xx, yy, xy) in 2D or
6 components (xx, yy, zz, xy, yz, xz) in 3D.
"""
import math

import numpy as np


def iso_ring_metric_2d(coords):
    """2D ring: h = 0.02 + 0.08 |r - 0.25|, r the distance to (0.5, 0.5)."""
    p = coords.reshape(-1, 2)
    r = np.sqrt((p[:, 0] - 0.5) ** 2 + (p[:, 1] - 0.5) ** 2)
    h = 0.02 + 0.08 * np.abs(r - 0.25)
    return 1.0 / (h * h), 1


def levelset_metric_3d(coords, h_min=0.002, h_max=0.05, width=0.02,
                       amplitude=0.1):
    """3D anisotropic metric around the level set phi = z - 0.5 - A sin(2 pi x).

    The normal eigen-length is h_n = h_min + (h_max - h_min)(1 - exp(-|phi|/w)),
    the two tangential eigen-lengths are h_max, and the eigenvectors follow
    grad(phi).  With two equal tangential eigenvalues the tensor is
    M = l_t I + (l_n - l_t) n n^T for l = 1/h^2.
    """
    p = coords.reshape(-1, 3)
    x, z = p[:, 0], p[:, 2]
    phi = z - 0.5 - amplitude * np.sin(2.0 * math.pi * x)
    g = np.stack([-2.0 * math.pi * amplitude * np.cos(2.0 * math.pi * x),
                  np.zeros_like(x), np.ones_like(x)], axis=1)
    n = g / np.linalg.norm(g, axis=1, keepdims=True)
    hn = h_min + (h_max - h_min) * (1.0 - np.exp(-np.abs(phi) / width))
    ln = 1.0 / (hn * hn)
    lt = 1.0 / (h_max * h_max)
    d = (ln - lt)
    m = np.empty((len(p), 6))
    m[:, 0] = lt + d * n[:, 0] * n[:, 0]
    m[:, 1] = lt + d * n[:, 1] * n[:, 1]
    m[:, 2] = lt + d * n[:, 2] * n[:, 2]
    m[:, 3] = d * n[:, 0] * n[:, 1]
    m[:, 4] = d * n[:, 1] * n[:, 2]
    m[:, 5] = d * n[:, 0] * n[:, 2]
    return m.ravel(), 6


def rotated_constant_metric(nverts, dim, rng):
    """Constant anisotropic metric: random rotation, eigen-lengths log-uniform
    in [0.005, 0.1]."""
    if dim == 3:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    else:
        a = rng.uniform(0, 2 * math.pi)
        R = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    h = np.exp(rng.uniform(math.log(0.005), math.log(0.1), dim))
    M = R @ np.diag(1.0 / (h * h)) @ R.T
    if dim == 3:
        v = np.array([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[1, 2], M[0, 2]])
    else:
        v = np.array([M[0, 0], M[1, 1], M[0, 1]])
    return np.tile(v, nverts), len(v)


def banded_aniso_metric(coords, dim, rng):
    """Anisotropic metric with a fixed random rotation whose two leading
    eigen-lengths grow away from the plane x = 0.5: h1 = 0.02 + 0.3|x - 0.5|,
    h2 = 2.5 h1, h3 = 0.15.  Steep enough near the plane for gradation
    limiting to change it over several sweeps."""
    q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
    p = coords.reshape(-1, dim)
    h1 = 0.02 + 0.3 * np.abs(p[:, 0] - 0.5)
    hs = np.stack([h1, 2.5 * h1, np.full_like(h1, 0.15)][:dim], axis=1)
    lam = 1.0 / (hs * hs)
    m = np.einsum("ij,nj,kj->nik", q, lam, q)
    if dim == 3:
        v = np.stack([m[:, 0, 0], m[:, 1, 1], m[:, 2, 2], m[:, 0, 1], m[:, 1, 2], m[:, 0, 2]], 1)
    else:
        v = np.stack([m[:, 0, 0], m[:, 1, 1], m[:, 0, 1]], 1)
    return v.ravel(), v.shape[1]


# the reference's empirical volume of an element of a unit-edge-length mesh
# (typical_unit_simplex_size, src/Omega_h_metric.cpp:286-290)
UNIT_SIMPLEX_SIZE = {2: 0.377645136635, 3: 0.0838934100219}


def symm_det(v):
    """Determinant of one metric in Omega_h's symmetric layout."""
    if len(v) == 3:
        return v[0] * v[1] - v[2] * v[2]
    a, b, c, d, e, f = v
    return a * (b * c - e * e) - d * (d * c - e * f) + f * (d * e - b * f)


def expected_nelems_unit_box(v, dim):
    """get_expected_nelems (src/Omega_h_metric.cpp:464-519) of a constant
    metric on the unit box: sqrt(det M) / the unit simplex's size."""
    return math.sqrt(symm_det(v)) / UNIT_SIMPLEX_SIZE[dim]


def clamp_metric(v, dim, max_nelems):
    """The reference's element-count limit (should_limit_element_count,
    src/Omega_h_metric_input.cpp:226-236 with get_metric_scalar_for_nelems,
    src/Omega_h_metric.cpp:521-526): a metric expected to give more than
    max_nelems elements is scaled by (max_nelems / expected)^(2/dim)."""
    n = expected_nelems_unit_box(v, dim)
    if n <= max_nelems:
        return np.asarray(v, dtype=np.float64), n, 1.0
    s = (max_nelems / n) ** (2.0 / dim)
    return np.asarray(v, dtype=np.float64) * s, n, s


def batch_boxes(nboxes, seed, n_range=(8, 48), h_range=(0.005, 0.1), max_nelems=None,
                dim=3):
    """Independent boxes with their own constant rotated metrics: cells per
    side n = round(exp(U(ln n0, ln n1))), a random unit quaternion (a
    normalised 4-D Gaussian) for the axes and eigen-lengths exp(U(ln h0, ln
    h1)); numpy's PCG64 generator seeded with `seed`.  Returns a list of
    (n, metric components, expected elements before the clamp, clamp scale)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nboxes):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        if dim == 3:
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            w, x, y, z = q
            R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                          [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                          [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        else:
            a = rng.uniform(0, 2 * math.pi)
            R = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
        h = np.exp(rng.uniform(math.log(h_range[0]), math.log(h_range[1]), dim))
        M = R @ np.diag(1.0 / (h * h)) @ R.T
        if dim == 3:
            v = [M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[1, 2], M[0, 2]]
        else:
            v = [M[0, 0], M[1, 1], M[0, 1]]
        if max_nelems is None:
            out.append((n, np.asarray(v), expected_nelems_unit_box(v, dim), 1.0))
        else:
            vc, e, s = clamp_metric(v, dim, max_nelems)
            out.append((n, vc, e, s))
    return out


def serve_request(rng, dim, elems, h_lo, h_hi, clamp=None):
    """One request of the serving workload: a box of about `elems` elements
    (2 n^2 triangles or 6 n^3 tets) under a constant anisotropic metric with
    a uniformly random rotation and eigen-lengths log-uniform in [h_lo, h_hi],
    scaled by the reference's element-count rule when it is expected to give
    more than `clamp` elements.  Returns (dim, n, metric components)."""
    n = max(1, round((elems / 2) ** 0.5)) if dim == 2 else max(1, round((elems / 6) ** (1 / 3)))
    g = rng.normal(size=(dim, dim))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))
    h = np.exp(rng.uniform(np.log(h_lo), np.log(h_hi), dim))
    m = q @ np.diag(1.0 / h ** 2) @ q.T
    if dim == 3:
        v = [m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[1, 2], m[0, 2]]
    else:
        v = [m[0, 0], m[1, 1], m[0, 1]]
    if clamp is not None:
        v = list(clamp_metric(v, dim, clamp)[0])
    return dim, n, v


def serve_stream(rng, rate, duration, max_elems, h_lo, h_hi, clamp=None, min_elems=1e3,
                 frac_2d=0.3):
    """Poisson arrivals at `rate` per second for `duration` seconds: a list of
    (arrival time, request) with 30 % 2D and 70 % 3D boxes and element
    counts log-uniform in [min_elems, max_elems]."""
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        dim = 2 if rng.random() < frac_2d else 3
        elems = float(np.exp(rng.uniform(np.log(min_elems), np.log(max_elems))))
        out.append((t, serve_request(rng, dim, elems, h_lo, h_hi, clamp)))
