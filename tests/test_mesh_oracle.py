"""Pins of the oracle's surface mesh (Contourer<3>::contour at 0,
Contourer.cpp:9-97, with SurfaceMesh::computeNormals, SurfaceMesh.cpp:
191-216):

* vertex numbering and positions against a brute-force restatement of the
  reference's creation rule on a small random field (cubes in order, each
  cube's edges 0..11, a vertex made the first time a crossing edge is met);
* a plane: every vertex lies on it, one vertex per crossing edge (counted
  in closed form);
* a sphere: closed and consistently oriented (every undirected edge in two
  triangles, every directed edge in one), Euler characteristic 2, enclosed
  volume 4/3 pi R^3 from the divergence theorem (positive: normals point
  out of the liquid), vertex normals radial;
* random fields: the only edges used by a single triangle lie on the grid's
  outer faces, and every other edge is used as often in each direction
  (the ambiguous faces are split the same way from both cubes; a polygon
  that crosses one face twice can put a fan chord on that face, where the
  two cubes' sheets then touch along it, four triangles on one edge);
* the table: each cube type's triangles use every crossing edge, and every
  triangle's normal points from the inside corners to the outside ones.
"""
import numpy as np
import pytest

import oracle as O


def _edge_corner(i, end):
    a = i >> 2
    p = [0, 0, 0]
    p[(a + 1) % 3] = i & 1
    p[(a + 2) % 3] = i >> 1 & 1
    p[a] = end
    return p


def _brute_vertices(f, o, dx):
    nz, ny, nx = f.shape
    seen, out = set(), []
    for k in range(nz - 1):
        for j in range(ny - 1):
            for i in range(nx - 1):
                for e in range(12):
                    a = e >> 2
                    p0 = np.array([i, j, k]) + np.array(_edge_corner(e, 0))
                    p1 = p0.copy()
                    p1[a] += 1
                    f0, f1 = f[p0[2], p0[1], p0[0]], f[p1[2], p1[1], p1[0]]
                    if (f0 <= 0) == (f1 <= 0):
                        continue
                    key = (a, *p0)
                    if key in seen:
                        continue
                    seen.add(key)
                    th = (f0 - 0) / (f0 - f1)
                    out.append([(1 - th) * (o[q] + p0[q] * dx) + th * (o[q] + p1[q] * dx) for q in range(3)])
    return np.array(out, np.float32)


def test_vertex_numbering_by_brute_force():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(5, 4, 6))
    o, dx = [0.1, -0.2, 0.3], 0.25
    pos, nrm, tri = O.contour3(f, o, dx)
    want = _brute_vertices(f, o, dx)
    assert pos.shape == want.shape and np.array_equal(pos, want)
    assert tri.max() < len(pos) and len(np.unique(tri)) == len(pos)


def _grid(n, dx=None):
    dx = dx or 1.0 / (n - 1)
    ax = np.arange(n) * dx
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return x, y, z, dx


def test_plane():
    x, y, z, dx = _grid(12)
    a, b, c, d = 0.3, -0.5, 0.8, -0.07
    f = a * x + b * y + c * z + d
    pos, _, tri = O.contour3(f, [0, 0, 0], dx)
    nrm = np.array([a, b, c]) / np.linalg.norm([a, b, c])
    assert np.max(np.abs(pos.astype(np.float64) @ np.array([a, b, c]) + d)) < 1e-6
    inside = f <= 0
    crossings = sum(int(np.sum(np.diff(inside.astype(np.int8), axis=ax) != 0)) for ax in range(3))
    assert len(pos) == crossings
    # every triangle faces +grad f
    p = pos.astype(np.float64)[tri]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    area = np.linalg.norm(fn, axis=1)
    assert np.all(fn[area > 1e-12] @ nrm > 0)


def _edges(tri):
    d = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    return d


def test_sphere_closed_oriented_volume_normals():
    x, y, z, dx = _grid(41)
    c, R = np.array([0.5, 0.48, 0.52]), 0.3
    f = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - R
    pos, nrm, tri = O.contour3(f, [0, 0, 0], dx)
    d = _edges(tri)
    und = np.sort(d, axis=1)
    _, cnt = np.unique(und, axis=0, return_counts=True)
    assert np.all(cnt == 2)
    _, dcnt = np.unique(d, axis=0, return_counts=True)
    assert np.all(dcnt == 1)
    V, E, F = len(pos), len(cnt), len(tri)
    assert V - E + F == 2
    p = pos.astype(np.float64)[tri]
    vol = np.sum(np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2]))) / 6
    assert vol == pytest.approx(4 / 3 * np.pi * R ** 3, rel=5e-3)
    radial = pos - c
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    assert np.min(np.einsum("ij,ij->i", radial, nrm)) > 0.98
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_field_open_edges_only_on_the_grid_boundary(seed):
    rng = np.random.default_rng(seed)
    f = rng.choice([-1.0, 1.0], size=(9, 8, 10)) * rng.uniform(0.5, 1.5, size=(9, 8, 10))
    pos, _, tri = O.contour3(f, [0, 0, 0], 1.0)
    d = _edges(tri)
    und = np.sort(d, axis=1)
    keys, inv, cnt = np.unique(und, axis=0, return_inverse=True, return_counts=True)
    # orientation: every undirected edge is crossed as often each way, except
    # the open edges on the grid boundary
    fwd = np.bincount(inv.ravel(), weights=(d[:, 0] < d[:, 1]).astype(float), minlength=len(keys))
    balanced = fwd * 2 == cnt
    assert np.all(balanced | (cnt == 1))
    lo, hi = np.zeros(3), np.array([9, 7, 8])   # x, y, z extents of the node grid
    for e in keys[cnt == 1]:
        p = pos[e].astype(np.float64)
        on = [(np.abs(p[:, q] - lo[q]) < 1e-9).all() or (np.abs(p[:, q] - hi[q]) < 1e-9).all() for q in range(3)]
        assert any(on)


def _polygons(row, n):
    """Fans back to polygons: consecutive triangles (r, a, b), (r, b, c), ..."""
    polys = []
    for q in range(n):
        r, a, b = row[3 * q:3 * q + 3]
        if polys and polys[-1][0] == r and polys[-1][-1] == a:
            polys[-1].append(b)
        else:
            polys.append([r, a, b])
    return polys


def test_table_uses_every_crossing_edge_and_points_out():
    """Every crossing edge is used, and each polygon's area vector (edge
    midpoints) has a positive component along the inside-to-outside
    directions of its edges -- an orientation rule independent of the face
    walk that built the table."""
    ntri, tri = O.mc_table()
    corner = lambda p: p[0] | p[1] << 1 | p[2] << 2
    for t in range(256):
        inside = [(t >> k & 1) == 1 for k in range(8)]
        cross = {i for i in range(12) if inside[corner(_edge_corner(i, 0))] != inside[corner(_edge_corner(i, 1))]}
        used = set(tri[t, :3 * ntri[t]].tolist())
        assert used == cross, t
        seen = []
        for poly in _polygons(tri[t], ntri[t]):
            assert poly[0] == min(poly) and len(set(poly)) == len(poly)
            seen += poly
            mid = [np.add(_edge_corner(i, 0), _edge_corner(i, 1)) / 2 for i in poly]
            N = sum(np.cross(mid[q], mid[(q + 1) % len(mid)]) for q in range(len(mid)))
            R = np.zeros(3)
            for i in poly:
                R[i >> 2] += 1 if inside[corner(_edge_corner(i, 0))] else -1
            assert N @ R > 0, (t, poly)
        assert sorted(seen) == sorted(cross)   # each crossing edge in one polygon
    assert ntri.max() <= 8


# ------------------------------------------------------------------ 2D ----
# Contourer<2, true>::contour, the filled contour of 2D frames.

def _brute_vertices2(f, o, dx):
    ny, nx = f.shape
    out = [[o[0] + i * dx, o[1] + j * dx] for j in range(ny) for i in range(nx) if f[j, i] <= 0]
    seen = set()
    for j in range(ny - 1):
        for i in range(nx - 1):
            for e in range(4):
                axis = e >> 1
                p0 = [i + (e & 1 if axis == 1 else 0), j + (e & 1 if axis == 0 else 0)]
                p1 = [p0[0] + (axis == 0), p0[1] + (axis == 1)]
                f0, f1 = f[p0[1], p0[0]], f[p1[1], p1[0]]
                if (f0 <= 0) == (f1 <= 0) or (axis, *p0) in seen:
                    continue
                seen.add((axis, *p0))
                th = (f0 - 0) / (f0 - f1)
                out.append([(1 - th) * (o[q] + p0[q] * dx) + th * (o[q] + p1[q] * dx) for q in range(2)])
    return np.array(out)


def test_filled_contour_vertex_numbering_by_brute_force():
    rng = np.random.default_rng(4)
    f = rng.normal(size=(7, 9))
    pos, tri = O.contour2(f, [0.1, -0.3], 0.5)
    assert np.array_equal(pos, _brute_vertices2(f, [0.1, -0.3], 0.5))
    assert tri.max() < len(pos)


def _signed_areas(pos, tri):
    p = pos[tri]
    return 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0]))


@pytest.mark.parametrize("seed", [0, 1])
def test_filled_contour_area_per_cell_and_orientation(seed):
    """Every triangle counter-clockwise, and the triangles of each cell
    cover the filled area of that cell (the 2D volume's per-cell area,
    itself pinned by half-planes and a disk) -- random signs, so every cell
    type and both diagonal rules occur."""
    rng = np.random.default_rng(seed)
    f = rng.choice([-1.0, 1.0], size=(12, 14)) * rng.uniform(0.2, 2.0, size=(12, 14))
    pos, tri = O.contour2(f, [0, 0], 1.0)
    area = _signed_areas(pos, tri)
    assert np.all(area > 0)
    want = sum(O.filled_area([f[j, i], f[j, i + 1], f[j + 1, i], f[j + 1, i + 1]]) for j in range(11) for i in range(13))
    assert area.sum() == pytest.approx(want, rel=1e-12)


def test_filled_disk():
    n = 81
    ax = np.arange(n) / (n - 1)
    y, x = np.meshgrid(ax, ax, indexing="ij")
    R = 0.3
    f = np.hypot(x - 0.5, y - 0.47) - R
    pos, tri = O.contour2(f, [0, 0], 1.0 / (n - 1))
    area = _signed_areas(pos, tri)
    assert np.all(area > 0)
    assert area.sum() == pytest.approx(np.pi * R * R, rel=2e-3)
    # a disk's triangulation: Euler characteristic 1, boundary edges once
    d = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    und, cnt = np.unique(np.sort(d, axis=1), axis=0, return_counts=True)
    assert np.all(cnt <= 2)
    assert len(pos) - len(und) + len(tri) == 1
