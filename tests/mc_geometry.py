"""Closed-form MC-style volumes of the ambiguous corner configurations, from
the geometry of the marching-cubes surface alone (no case table).

Corners are numbered in the fraction's circular order (fraction.hpp:44-51).
Two corners are adjacent when they share a cube edge.  For the cube values
used here -- inside corners at -alpha_i, outside corners at +1 -- the
marching-cubes surface around a set of inside corners that touch no other
inside corner is:

* an isolated corner: one triangle through the three edge crossings, each at
  distance l = alpha / (1 + alpha) from the corner; the inside region is a
  right tetrahedron, volume l^3 / 6;
* an isolated edge with corner values -alpha, -beta: a quadrilateral through
  the crossings at a = alpha / (1 + alpha) on both edges leaving the first
  corner and b = beta / (1 + beta) on both leaving the second.  Those four
  points are coplanar (a * b = b * a), so the region is
  {s in [0, 1], y, z >= 0, y + z <= a + (b - a) s}, volume (a^2 + ab + b^2) / 6.

The MC-style evaluation (Takahashi and Batty, JCGT 11(2), 2022) keeps such
components apart, so the volume of the configuration is the sum over its
components.  The same holds with inside and outside exchanged for two
outside corners on a body diagonal (value +beta each, inside corners at -1):
volume 1 - sum of the two outside tetrahedra.
"""
import itertools

import numpy as np

CIRC = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
_P = np.array(CIRC)
ADJ = {i: [j for j in range(8) if int(np.abs(_P[i] - _P[j]).sum()) == 1] for i in range(8)}


def components(corners):
    """Edge-connected components of a set of corners."""
    left, out = set(corners), []
    while left:
        seed = left.pop()
        comp, stack = {seed}, [seed]
        while stack:
            x = stack.pop()
            for y in ADJ[x]:
                if y in left:
                    left.remove(y)
                    comp.add(y)
                    stack.append(y)
        out.append(sorted(comp))
    return out


def inside_set(pattern):
    return [i for i in range(8) if pattern >> i & 1]


def corners_and_edges(pattern):
    """True when the inside corners (at most four) form isolated corners and
    isolated edges only."""
    ins = inside_set(pattern)
    return 0 < len(ins) <= 4 and all(len(c) <= 2 for c in components(ins))


def separated_volume(pattern, alpha):
    """Volume of a corners_and_edges pattern with inside corner i at -alpha[i]
    and every outside corner at +1."""
    vol = 0.0
    for comp in components(inside_set(pattern)):
        if len(comp) == 1:
            l = alpha[comp[0]] / (1.0 + alpha[comp[0]])
            vol += l ** 3 / 6.0
        else:
            a = alpha[comp[0]] / (1.0 + alpha[comp[0]])
            b = alpha[comp[1]] / (1.0 + alpha[comp[1]])
            vol += (a * a + a * b + b * b) / 6.0
    return vol


BODY_DIAGONALS = [(0, 6), (1, 7), (2, 4), (3, 5)]


def body_diagonal_complement_volume(beta_a, beta_b):
    """Volume when two outside corners on a body diagonal sit at +beta_a,
    +beta_b and the six others at -1."""
    la, lb = beta_a / (1.0 + beta_a), beta_b / (1.0 + beta_b)
    return 1.0 - (la ** 3 + lb ** 3) / 6.0


def rotations():
    """The 24 proper rotations of the cube as corner permutations: corner i
    moves to perm[i]."""
    perms = []
    for axes in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for r in range(3):
                m[r, axes[r]] = signs[r]
            if np.linalg.det(m) < 0:
                continue
            q = (_P - 0.5) @ m.T + 0.5
            perms.append([int(np.argmin(np.abs(_P - q[i]).sum(axis=1))) for i in range(8)])
    return perms


def orbit_class(pattern):
    """Rotation-invariant description of a pattern: inside count, component
    sizes of the inside and the outside corners, and, for two isolated
    corners, whether they share a face."""
    ins = inside_set(pattern)
    outs = [i for i in range(8) if i not in ins]
    ci, co = components(ins), components(outs)

    def pair_kind(cs):
        if len(cs) == 2 and all(len(c) == 1 for c in cs):
            return int(np.abs(_P[cs[0][0]] - _P[cs[1][0]]).sum())   # 2: face diagonal, 3: body diagonal
        return 0
    return (len(ins), tuple(sorted(len(c) for c in ci)), tuple(sorted(len(c) for c in co)), pair_kind(ci), pair_kind(co))


# ---------------------------------------------------------------------------
# Candidate volumes of any sign pattern from the marching-cubes geometry.
#
# The inside region of a cube under a marching-cubes surface is bounded by
# the inside parts of the six cube faces and by the surface's polygons, whose
# vertices are the linear zero crossings on the cube edges.  On a face with
# two inside corners on a diagonal the crossings pair up either around the
# inside corners or around the outside ones; every other face has one
# pairing.  Once each face is resolved, the segments close into loops, and a
# loop with more than three vertices is generally skew: its triangulation
# (without extra vertices) changes the volume.  By the divergence theorem with
# the field (x, 0, 0),
#     V = area of the inside part of the face x = 1
#       + sum over surface triangles of  centroid_x * (outward area vector)_x,
# so every admissible surface gives one candidate volume.  Any MC-style table
# must return one of them; a dropped term, a wrong sign or a wrong corner in a
# case formula lands on none.

EDGES = sorted((i, j) for i in range(8) for j in ADJ[i] if i < j)


def _faces():
    out = []
    for a in range(3):
        for s in (0, 1):
            cs = [i for i in range(8) if CIRC[i][a] == s]
            f = np.zeros(3)
            f[a] = 2 * s - 1
            u = np.zeros(3)
            u[(a + 1) % 3] = 1.0
            w = np.cross(f, u)
            ctr = _P[cs].mean(axis=0)
            ang = [np.arctan2(np.dot(_P[i] - ctr, w), np.dot(_P[i] - ctr, u)) for i in cs]
            cs = [c for _, c in sorted(zip(ang, cs))]   # counter-clockwise seen from outside
            out.append((a, s, f, u, w, cs))
    return out


FACES = _faces()


def _crossing(vals, e):
    i, j = e
    t = vals[i] / (vals[i] - vals[j])
    return _P[i] + t * (_P[j] - _P[i])


def _edge(i, j):
    return (i, j) if i < j else (j, i)


def _face_resolutions(vals, face):
    """Every admissible set of segments on one face: (edge from, edge to,
    corners cut off, in boundary order)."""
    cs = face[5]
    ins = [vals[c] < 0 for c in cs]
    if sum(ins) in (0, 4):
        return [[]]
    cross = [k for k in range(4) if ins[k] != ins[(k + 1) % 4]]
    if len(cross) == 2:
        k1, k2 = cross
        cut = [cs[m % 4] for m in range(k1 + 1, k2 + 1)]
        return [[(_edge(cs[k1], cs[(k1 + 1) % 4]), _edge(cs[k2], cs[(k2 + 1) % 4]), cut)]]
    opts = []
    for want in (True, False):   # cut off the inside corners, or the outside ones
        opts.append([(_edge(cs[(m - 1) % 4], cs[m]), _edge(cs[m], cs[(m + 1) % 4]), [cs[m]]) for m in range(4) if ins[m] == want])
    return opts


def _piece_area(vals, face, seg):
    """Area of the face piece a segment cuts off (its corners in boundary order)."""
    u, w = face[3], face[4]
    pts = [_crossing(vals, seg[0])] + [_P[c].astype(float) for c in seg[2]] + [_crossing(vals, seg[1])]
    xy = np.array([[np.dot(p, u), np.dot(p, w)] for p in pts])
    x, y = xy[:, 0], xy[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _tri_sums(pts):
    """Sum of centroid_x * area-vector_x over every triangulation of a loop."""
    n = len(pts)
    memo = {}

    def term(i, k, j):
        a, b, c = pts[i], pts[k], pts[j]
        return (a[0] + b[0] + c[0]) / 3.0 * 0.5 * np.cross(b - a, c - a)[0]

    def sums(i, j):
        if j - i < 2:
            return np.zeros(1)
        if (i, j) not in memo:
            memo[(i, j)] = np.concatenate([term(i, k, j) + np.add.outer(sums(i, k), sums(k, j)).ravel() for k in range(i + 1, j)])
        return memo[(i, j)]
    return sums(0, n - 1)


def candidate_volumes(vals):
    """Volumes of the inside region {v < 0} for every face resolution and
    every triangulation of the resulting loops (vals in CIRC order)."""
    vals = np.asarray(vals, dtype=np.float64)
    per_face = [_face_resolutions(vals, f) for f in FACES]
    out = []
    for choice in itertools.product(*per_face):
        directed = {}
        area1 = 0.0
        for face, segs in zip(FACES, choice):
            outside_pieces = []
            for seg in segs:
                p, q = _crossing(vals, seg[0]), _crossing(vals, seg[1])
                c = seg[2][0]
                side = np.dot(np.cross(q - p, _P[c] - p), face[2])
                # walking the loop, the face's inside lies on the right (seen from outside)
                ein, eout = (seg[0], seg[1]) if (side < 0) == (vals[c] < 0) else (seg[1], seg[0])
                assert ein not in directed, "inconsistent loop orientation"
                directed[ein] = eout
                if face[0] == 0 and face[1] == 1:
                    if vals[c] < 0:
                        area1 += _piece_area(vals, face, seg)
                    else:
                        outside_pieces.append(_piece_area(vals, face, seg))
            if face[0] == 0 and face[1] == 1:
                if outside_pieces:
                    area1 += 1.0 - sum(outside_pieces)
                elif not segs and vals[face[5][0]] < 0:
                    area1 += 1.0
        total = np.array([area1])
        left = dict(directed)
        while left:
            e0 = next(iter(left))
            loop, e = [], e0
            while True:
                loop.append(_crossing(vals, e))
                e = left.pop(e)
                if e == e0:
                    break
            total = np.add.outer(total, _tri_sums(loop)).ravel()
        out.append(total)
    return np.concatenate(out)
