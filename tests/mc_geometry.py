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
