"""Grid convention and scenes of the benchmark configurations.

Grid: n array cells per axis including a solid band of `bw` cells on each
side, as the reference's builder lays out StaggeredGrid(bw, L/(n - 2 bw), n)
(LevelSetLiquidBuilder.cpp:49,74); the origin is chosen so the liquid domain
is [0, 1]^D (domainOrigin = 0, StaggeredGrid.h:40).

A scene acts on any object with union_box / union_sphere / union_plane /
get / set / count (the device Liquid, or the CPU oracle in tests).
"""
import numpy as np


def grid(n, dim, bw=2):
    dx = 1.0 / (n - 2 * bw)
    return dict(dim=dim, res=[n] * dim, dx=dx, origin=[-bw * dx] * dim, bw=bw)


def face_positions(g, axis):
    """Positions of the faces of `axis` (StaggeredGrid::makeFaceGrids,
    StaggeredGrid.h:104-120), x-fastest, shape (count, dim)."""
    dim, dx = g["dim"], g["dx"]
    n = [r + (1 if a == axis else 0) for a, r in enumerate(g["res"])]
    o = [g["origin"][a] + (0.0 if a == axis else dx / 2) for a in range(dim)]
    axes = [o[a] + np.arange(n[a]) * dx for a in range(dim)]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    return np.stack([m.ravel() for m in mesh], axis=1)


def cell_positions(g):
    dim, dx = g["dim"], g["dx"]
    axes = [g["origin"][a] + dx / 2 + np.arange(g["res"][a]) * dx for a in range(dim)]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    return np.stack([m.ravel() for m in mesh], axis=1)


def dam_break(sim, g, extents=(0.4, 0.6, 0.5)):
    """Liquid box x < e0, y < e1 (, z < e2) anchored at the domain origin,
    as buildCase1 anchors its box (LevelSetLiquidBuilder.cpp:79)."""
    dim = g["dim"]
    sim.union_box([0.0] * dim, list(extents[:dim]))


def droplet(sim, g, seed=0, radius=0.08, pool=0.3, height=0.7, speed=-2.0):
    """Pool phi = y - pool union a sphere at (x, height, z), x, z ~ U[0.3, 0.7]
    (seeded), the drop moving at `speed` along y (buildCase0 pattern,
    LevelSetLiquidBuilder.cpp:57-62)."""
    dim = g["dim"]
    rng = np.random.default_rng(seed)
    c = [float(rng.uniform(0.3, 0.7)), height] + ([float(rng.uniform(0.3, 0.7))] if dim == 3 else [])
    sim.union_plane([0.0, pool, 0.0][:dim], [0.0, 1.0, 0.0][:dim])
    sim.union_sphere(c, radius)
    u = np.zeros(sim.count(9))
    off = 0
    for a in range(dim):
        pos = face_positions(g, a)
        if a == 1:
            inside = np.linalg.norm(pos - np.array(c)[None, :], axis=1) < radius
            u[off:off + len(pos)][inside] = speed
        off += len(pos)
    sim.set(9, u)
    return c


def dam_break_with_drop(sim, g, extents, drop, radius=0.08):
    """A dam-break box (extents from the origin) and a spherical drop of
    `radius` at `drop`, both at rest: the scenes of the network's training
    runs (box extents and drop positions drawn by the caller)."""
    dim = g["dim"]
    sim.union_box([0.0] * dim, list(extents[:dim]))
    sim.union_sphere(list(drop[:dim]), radius)


def capillary_grid(scale, dim, bw=2):
    """The grid of LevelSetLiquidBuilder::buildCase1 (LevelSetLiquidBuilder.cpp:67-76):
    resolution scale (1, 1, 1) + scale e0 + int(0.2 scale) e1, spacing
    1 / (scale - 2 bw), the domain's lower corner at the origin."""
    dx = 1.0 / (scale - 2 * bw)
    res = [2 * scale, scale + int(0.2 * scale), scale][:dim]
    return dict(dim=dim, res=res, dx=dx, origin=[-bw * dx] * dim, bw=bw)


def capillary_column(sim, g):
    """buildCase1's liquid (LevelSetLiquidBuilder.cpp:79): the box from the
    domain origin with lengths (0.4, 0.4, 1).  Its solid cylinder
    (:82, base (0, -0.6, 0), radius 0.1, height 0.6) lies in y <= 0, below
    the domain floor, where the domain walls are solid already: it changes no
    node distance above the floor, so the domain box alone is its boundary."""
    dim = g["dim"]
    sim.union_box([0.0] * dim, [0.4, 0.4, 1.0][:dim])


# buildCase1's surface-tension settings (LevelSetLiquidBuilder.cpp:76-78):
# _surfaceTensionCoeff 30 over the default density 1e3, band 5.5 cells
CAPILLARY_COEFF = 30.0 / 1e3
CAPILLARY_BAND = 5.5
