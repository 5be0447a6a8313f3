"""Pins of the CPU oracle by what the mathematics fixes.

The reference ships no golden vectors (its tests print metrics only), so each
restated operator is checked against closed forms, exactness properties and
hand-derived values chosen so that a plausible slip (a dropped term, a wrong
sign, index or transposition) fails:

* boundary: cell distance to the domain walls, face fractions of a box domain;
* GARM advection: affine maps under uniform and linear velocity (closed-form
  RK2 backtrace and Jacobian), polynomial reference maps reproduced by the
  tensor Hermite interpolant with its finite-difference mixed derivatives;
* cubic velocity advection: linear velocity advected to its closed form;
* composition: reproduction of a cubic reference level set at arbitrary
  reference-map positions, grad phi = J^T grad phi_ref;
* WENO5 reinit: a planar signed distance is a fixed point, a distorted
  circle converges to the true distance (TestReinit.cpp:89-171 set-up);
* projection: a single liquid cell solved by hand (theta = 1/2 and the
  1e-3 clamp), discrete divergence after projection equals the solver
  residual, three preconditioners reach the same solution;
* extrapolation: a single seed spreads as an L1 ball of the stated radius;
* boundary enforcement: inflow removed, 0.2 of the tangential slip kept;
* fourth-order gradient (Derivatives::computeFirst<4>, taken at restart and
  initialisation): exact for polynomials of degree 4 per axis away from the
  walls, and the clamped edge stencil by hand on a 6 x 5 grid.
"""
import numpy as np
import pytest

import oracle as O
from garm_amd import scenes

PHI, PHI_GRAD, REFMAP, REFMAP_GRAD, PHI_TMP, PHI_ORI, PHI_REF, PHI_REF_GRAD, ADV, VEL, CDIST, NDIST, FRAC, BVEL, NORMAL, PRES = range(16)


def mk(n, dim):
    g = scenes.grid(n, dim)
    return g, O.Oracle(dim, g["res"], g["dx"], g["origin"], g["bw"])


def face_split(g, arr):
    out, off = [], 0
    for a in range(g["dim"]):
        m = len(scenes.face_positions(g, a))
        out.append(arr[off:off + m])
        off += m
    return out


def interior_mask(g, margin):
    idx = np.indices(g["res"][::-1]).reshape(g["dim"], -1)[::-1]   # x-fastest coordinates
    m = np.ones(idx.shape[1], bool)
    for a in range(g["dim"]):
        m &= (idx[a] >= margin) & (idx[a] < g["res"][a] - margin)
    return m


# ------------------------------------------------------------ boundary ----
@pytest.mark.parametrize("dim", [2, 3])
def test_boundary_box_domain(dim):
    g, o = mk(12, dim)
    x = scenes.cell_positions(g)
    inside = np.all((x > 0) & (x < 1), axis=1)
    cd = o.get(CDIST)
    # inside a box the signed distance is the distance to the nearest wall
    expect = np.min(np.minimum(x, 1 - x), axis=1)
    assert np.allclose(cd[inside], expect[inside], atol=1e-14, rtol=0)
    assert np.all(cd[~inside] < 0)
    fr = face_split(g, o.get(FRAC))
    bw = g["bw"]
    for a in range(dim):
        pos = scenes.face_positions(g, a)
        n = [r + (1 if b == a else 0) for b, r in enumerate(g["res"])]
        idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]
        boundary = (idx[a] <= bw) | (idx[a] >= g["res"][a] - bw)
        for b in range(dim):
            boundary |= (idx[b] < bw) | (idx[b] >= n[b] - bw)
        assert np.all(fr[a][boundary] == 1.0)
        assert np.all(fr[a][~boundary] == 0.0), "a box domain has no partially solid interior face"
        assert len(pos) == len(fr[a])


# ------------------------------------------------------ GARM advection ----
def _set_affine_refmap(o, g):
    x = scenes.cell_positions(g)
    o.set(REFMAP, x.ravel())
    o.set(REFMAP_GRAD, np.tile(np.eye(g["dim"]).ravel(order="F"), len(x)))
    return x


def _set_linear_velocity(o, g, A):
    u = []
    for a in range(g["dim"]):
        p = scenes.face_positions(g, a)
        u.append(p @ A[a])
    o.set(VEL, np.concatenate(u))


@pytest.mark.parametrize("dim", [2, 3])
def test_refmap_uniform_translation(dim):
    g, o = mk(16, dim)
    x = _set_affine_refmap(o, g)
    c = np.array([0.7, -0.4, 0.3][:dim])
    o.set(VEL, np.concatenate([np.full(len(scenes.face_positions(g, a)), c[a]) for a in range(dim)]))
    dt = 0.6 * g["dx"]
    o.advect_refmap(dt)
    m = interior_mask(g, 4)
    xi = o.get(REFMAP).reshape(-1, dim)
    J = o.get(REFMAP_GRAD).reshape(-1, dim, dim)
    assert np.abs(xi[m] - (x[m] - c * dt)).max() < 1e-13
    assert np.abs(J[m] - np.eye(dim)).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_refmap_linear_velocity_closed_form(dim):
    """u = A x: RK2 backtrace X = x - dt A x + dt^2 A^2 x / 2 and
    R = I - dt A + dt^2 A^2 / 2 exactly (Advection.h:71-82), and an affine
    reference map is reproduced by the Hermite interpolant, so xi' = X, J' = R."""
    g, o = mk(20, dim)
    x = _set_affine_refmap(o, g)
    A = np.array([[0.3, -1.1, 0.2], [0.9, -0.2, 0.5], [-0.4, 0.6, 0.1]])[:dim, :dim]
    _set_linear_velocity(o, g, A)
    dt = 0.02
    o.advect_refmap(dt)
    h = -dt
    X = x + h * x @ A.T + h * h / 2 * x @ (A @ A).T
    R = np.eye(dim) + h * A + h * h / 2 * A @ A
    m = interior_mask(g, 5)
    xi = o.get(REFMAP).reshape(-1, dim)
    J = o.get(REFMAP_GRAD).reshape(-1, dim, dim).transpose(0, 2, 1)  # column-major storage
    assert np.abs(xi[m] - X[m]).max() < 1e-13
    assert np.abs(J[m] - R).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_refmap_hermite_reproduces_cubic_map(dim):
    """A reference map whose components are polynomials of total degree <= 3
    (with exact Jacobian data) is reproduced at backtraced points: the
    tensor-cubic Hermite space contains them and the finite-difference mixed
    derivatives (HermiteIntrpl.h:17-30) are exact for them."""
    g, o = mk(20, dim)
    x = scenes.cell_positions(g)
    rng = np.random.default_rng(1)

    def poly(p):
        # per component: random cubic of total degree <= 3
        out = np.zeros((len(p), dim))
        grad = np.zeros((len(p), dim, dim))
        for r in range(dim):
            cs = rng.normal(size=(dim, dim, dim))
            lin = rng.normal(size=dim)
            for i in range(dim):
                out[:, r] += lin[i] * p[:, i]
                grad[:, r, i] += lin[i]
                for j in range(dim):
                    out[:, r] += cs[i, j, 0] * p[:, i] * p[:, j] * p[:, (i + j) % dim] / 3
                    # d/dp_k of p_i p_j p_l
                    l = (i + j) % dim
                    for k in range(dim):
                        d = (p[:, j] * p[:, l] if i == k else 0) + (p[:, i] * p[:, l] if j == k else 0) + (p[:, i] * p[:, j] if l == k else 0)
                        grad[:, r, k] += cs[i, j, 0] * d / 3
        return out, grad

    state = rng.bit_generator.state
    val, grad = poly(x)
    o.set(REFMAP, val.ravel())
    o.set(REFMAP_GRAD, grad.transpose(0, 2, 1).ravel())   # store column-major
    c = np.array([0.45, -0.3, 0.25][:dim])
    o.set(VEL, np.concatenate([np.full(len(scenes.face_positions(g, a)), c[a]) for a in range(dim)]))
    dt = 0.7 * g["dx"]
    o.advect_refmap(dt)
    rng.bit_generator.state = state
    ev, eg = poly(x - c * dt)
    m = interior_mask(g, 4)
    xi = o.get(REFMAP).reshape(-1, dim)
    J = o.get(REFMAP_GRAD).reshape(-1, dim, dim).transpose(0, 2, 1)
    assert np.abs(xi[m] - ev[m]).max() < 1e-12
    assert np.abs(J[m] - eg[m]).max() < 1e-11


@pytest.mark.parametrize("dim", [2, 3])
def test_velocity_advection_linear(dim):
    """u = A x advects to (A X)_a at each face, X the RK2 backtrace of the
    face centre (Advection.h:37-43) — cubic Lagrange reproduces linear data."""
    g, o = mk(18, dim)
    A = np.array([[0.2, -0.7, 0.3], [0.5, -0.1, 0.4], [-0.3, 0.2, 0.05]])[:dim, :dim]
    _set_linear_velocity(o, g, A)
    dt = 0.03
    o.advect_velocity(dt)
    u = face_split(g, o.get(VEL))
    h = -dt
    for a in range(dim):
        p = scenes.face_positions(g, a)
        X = p + h * p @ A.T + h * h / 2 * p @ (A @ A).T
        n = [r + (1 if b == a else 0) for b, r in enumerate(g["res"])]
        idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]
        m = np.all([(idx[b] >= 4) & (idx[b] < n[b] - 4) for b in range(dim)], axis=0)
        assert np.abs(u[a][m] - (X @ A[a])[m]).max() < 1e-13


# ------------------------------------------------------------ compose ----
@pytest.mark.parametrize("dim", [2, 3])
def test_compose_reproduces_cubic_levelset(dim):
    g, o = mk(18, dim)
    x = scenes.cell_positions(g)
    c3 = np.array([0.7, -0.5, 0.3][:dim])
    c2 = np.array([[0.4, 0.1, -0.2], [0.1, -0.3, 0.2], [-0.2, 0.2, 0.5]])[:dim, :dim]

    def P(p):
        return 0.2 + p @ c3 + np.einsum("ni,ij,nj->n", p, c2, p) + 0.8 * p[:, 0] ** 2 * p[:, 1] - 0.5 * p[:, 0] * p[:, 1] * p[:, -1]

    def dP(p):
        d = c3[None, :] + 2 * p @ c2
        d[:, 0] += 1.6 * p[:, 0] * p[:, 1]
        d[:, 1] += 0.8 * p[:, 0] ** 2
        t = -0.5 * p[:, 0] * p[:, 1] * p[:, -1]
        # derivative of -0.5 x y w with w the last coordinate (w = y in 2D)
        if dim == 3:
            d[:, 0] += -0.5 * p[:, 1] * p[:, 2]
            d[:, 1] += -0.5 * p[:, 0] * p[:, 2]
            d[:, 2] += -0.5 * p[:, 0] * p[:, 1]
        else:
            d[:, 0] += -0.5 * p[:, 1] ** 2
            d[:, 1] += -1.0 * p[:, 0] * p[:, 1]
        del t
        return d

    o.set(PHI_REF, P(x))
    o.set(PHI_REF_GRAD, dP(x).ravel())
    rng = np.random.default_rng(5)
    xi = x + rng.uniform(-0.9, 0.9, size=x.shape) * g["dx"]
    B = np.eye(dim) + 0.1 * rng.normal(size=(dim, dim))
    o.set(REFMAP, xi.ravel())
    o.set(REFMAP_GRAD, np.tile(B.ravel(order="F"), len(x)))
    o.set(ADV, np.ones(len(x), np.uint8))
    o.compose()
    m = interior_mask(g, 4) & np.all((x > 0.05) & (x < 0.95), axis=1)
    phi = o.get(PHI)
    expect = np.maximum(P(xi), -o.get(CDIST) - 3 * g["dx"])
    assert np.abs(phi[m] - expect[m]).max() < 1e-12
    grad = o.get(PHI_GRAD).reshape(-1, dim)
    assert np.abs(grad[m] - dP(xi)[m] @ B).max() < 1e-11
    assert np.array_equal(o.get(PHI_ORI), phi)
    assert np.array_equal(o.get(PHI_TMP), np.maximum(phi, -o.get(CDIST)))


def test_compose_air_clamp():
    g, o = mk(12, 2)
    x = scenes.cell_positions(g)
    o.set(PHI_REF, np.full(len(x), -1.0))
    o.set(PHI_REF_GRAD, np.zeros(2 * len(x)))
    o.set(REFMAP, x.ravel())
    o.set(REFMAP_GRAD, np.tile(np.eye(2).ravel(), len(x)))
    adv = np.ones(len(x), np.uint8)
    adv[::3] = 0
    o.set(ADV, adv)
    o.compose()
    phi = o.get(PHI)
    floor = -o.get(CDIST) - 3 * g["dx"]
    assert np.allclose(phi[adv == 0], np.maximum(15 * g["dx"], floor[adv == 0]))
    assert np.allclose(phi[adv == 1], np.maximum(-1.0, floor[adv == 1]))


# ------------------------------------------------------------- reinit ----
@pytest.mark.parametrize("dim", [2, 3])
def test_reinit_plane_is_fixed_point(dim):
    g, o = mk(20, dim)
    x = scenes.cell_positions(g)
    phi = (x[:, 0] - 0.43) * 1.0
    o.set(PHI, phi)
    o.reinit(5, 0.9)
    m = interior_mask(g, 4)
    assert np.abs(o.get(PHI)[m] - phi[m]).max() < 1e-13


def test_reinit_distorted_circle_converges():
    """TestReinit.cpp:89-171: a circle's distance scaled by a smooth positive
    factor relaxes to the true distance near the interface."""
    g, o = mk(68, 2)
    x = scenes.cell_positions(g)
    r = np.linalg.norm(x - 0.5, axis=1)
    true = r - 0.25
    phi0 = true * (0.1 + (x[:, 0] - 0.3) ** 2 + (x[:, 1] - 0.6) ** 2)
    o.set(PHI, phi0)
    band = np.abs(true) < 3 * g["dx"]
    e0 = np.abs(phi0[band] - true[band]).mean()
    errs = []
    for _ in range(4):
        o.reinit(15, 0.9)
        errs.append(np.abs(o.get(PHI)[band] - true[band]).mean())
    # 15 iterations remove the distortion (about 45x); afterwards the PDE
    # only lets the interface creep (a few thousandths of a cell per 15)
    assert errs[0] < e0 / 30
    assert max(errs) < 0.05 * g["dx"]
    assert errs[-1] - errs[0] < 0.01 * g["dx"]


# ----------------------------------------------------------- projection ----
def test_projection_single_cell_by_hand():
    g, o = mk(10, 2)
    N = 100
    phi = np.ones(N)
    cid = 4 + 10 * 5
    phi[cid] = -1.0                    # theta = -1 / (-1 - 1) = 1/2 on every face
    o.set(PHI, phi)
    u = face_split(g, np.zeros(o.count(VEL)))
    u[0][5 + 11 * 5] = 1.0             # +x face of the cell (face index (5, 5) on an 11 x 10 grid)
    o.set(VEL, np.concatenate(u))
    r = o.project(precond=O.PRECOND_JACOBI, tol=1e-14)
    assert r["unknowns"] == 1
    p = o.get(PRES)
    assert p[cid] == pytest.approx(1 / 8, abs=1e-15)        # diag = 4 faces * w / theta = 8, rhs = +1
    ux, uy = face_split(g, o.get(VEL))
    assert ux[5 + 11 * 5] == pytest.approx(0.75, abs=1e-15)  # 1 + (0 - 1/8) * 2
    assert ux[4 + 11 * 5] == pytest.approx(0.25, abs=1e-15)  # (1/8 - 0) * 2
    assert uy[4 + 10 * 5] == pytest.approx(0.25, abs=1e-15)
    assert uy[4 + 10 * 6] == pytest.approx(-0.25, abs=1e-15)


def test_projection_interface_clamp():
    g, o = mk(10, 2)
    phi = np.full(100, 1e12)
    cid = 4 + 10 * 5
    phi[cid] = -1.0                    # theta ~ 1e-12 -> clamped to 1e-3
    o.set(PHI, phi)
    u = face_split(g, np.zeros(o.count(VEL)))
    u[1][4 + 10 * 6] = 2.0             # +y face
    o.set(VEL, np.concatenate(u))
    o.project(precond=O.PRECOND_IC0, tol=1e-14)
    assert o.get(PRES)[cid] == pytest.approx(2.0 / 4000.0, rel=1e-14)


def _divergence(g, o):
    """Discrete divergence of each liquid cell (Projector.cpp:30-68 rhs)."""
    dim = g["dim"]
    u = face_split(g, o.get(VEL))
    fr = face_split(g, o.get(FRAC))
    res = g["res"]
    div = np.zeros(res[::-1])
    for a in range(dim):
        n = [r + (1 if b == a else 0) for b, r in enumerate(res)]
        w = (1 - fr[a]).reshape(n[::-1])
        uu = u[a].reshape(n[::-1])
        flux = w * uu
        sl_hi = [slice(None)] * dim
        sl_lo = [slice(None)] * dim
        ax = dim - 1 - a
        sl_hi[ax] = slice(1, None)
        sl_lo[ax] = slice(0, -1)
        div += flux[tuple(sl_hi)] - flux[tuple(sl_lo)]
    return div.ravel()


@pytest.mark.parametrize("dim,n", [(2, 40), (3, 16)])
def test_projection_divergence_free_and_solver_independent(dim, n):
    g = scenes.grid(n, dim)
    sols = {}
    for pc in (O.PRECOND_NONE, O.PRECOND_JACOBI, O.PRECOND_IC0):
        o = O.Oracle(dim, g["res"], g["dx"], g["origin"], g["bw"])
        scenes.dam_break(o, g)
        o.initialize(tol=1e-12)
        o.gravity(0.04)
        o.advect_velocity(0.02)
        before = np.abs(_divergence(g, o)).max()
        r = o.project(precond=pc, tol=1e-13, max_it=100000)
        assert r["rc"] == 0
        phi = o.get(PHI)
        div = _divergence(g, o)
        liquid = np.abs(o.get(PRES)) > 0
        assert np.abs(div[liquid]).max() < 1e-9 * max(before, 1.0)
        sols[pc] = o.get(PRES)
        del phi
    ref = sols[O.PRECOND_IC0]
    for pc in (O.PRECOND_NONE, O.PRECOND_JACOBI):
        assert np.abs(sols[pc] - ref).max() <= 1e-8 * np.abs(ref).max()


# -------------------------------------------------------- extrapolation ----
def test_extrapolation_single_seed_is_l1_ball():
    g, o = mk(40, 2)
    N = 1600
    phi_ori = np.ones(N)
    seed = 20 + 40 * 19
    phi_ori[seed] = -1.0
    o.set(PHI_ORI, phi_ori)
    o.set(ADV, np.zeros(N, np.uint8))
    o.set(VEL, np.zeros(o.count(VEL)))
    o.extrapolate_air()
    adv = o.get(ADV).reshape(40, 40)
    jj, ii = np.indices((40, 40))
    ball = (np.abs(ii - 20) + np.abs(jj - 19)) <= 15
    assert np.array_equal(adv.astype(bool), ball)


def test_extrapolation_refmap_values_and_reset():
    g, o = mk(30, 2)
    N = 900
    x = scenes.cell_positions(g)
    adv = np.zeros(N, np.uint8)
    seed = 10 + 30 * 12
    adv[seed] = 1
    xi = x.copy()
    xi[seed] = [0.5, 0.25]
    o.set(ADV, adv)
    o.set(REFMAP, xi.ravel())
    o.set(PHI_ORI, np.ones(N))
    o.extrapolate_air()
    out = o.get(REFMAP).reshape(-1, 2)
    jj, ii = np.indices((30, 30))
    ball = ((np.abs(ii - 10) + np.abs(jj - 12)) <= 15).ravel()
    assert np.array_equal(out[ball], np.tile([0.5, 0.25], (ball.sum(), 1)))   # means of equal powers of two are exact
    assert np.array_equal(out[~ball], x[~ball])


def test_extrapolation_collider_fills_solid_band():
    g, o = mk(16, 3)
    N = 16 ** 3
    cd = o.get(CDIST)
    v = np.where(cd > 0, 0.5, np.nan)
    o.set(PHI, v)
    o.set(PHI_ORI, v)
    o.extrapolate_collider()
    # every solid-band cell lies within 6 L1 steps of the interior here
    assert np.all(o.get(PHI) == 0.5) and np.all(o.get(PHI_ORI) == 0.5)
    assert N == len(cd)


# ------------------------------------------------------------- enforce ----
def test_enforce_boundary_removes_inflow():
    g, o = mk(16, 2)
    ux = np.full(17 * 16, -1.0)
    uy = np.full(16 * 17, 0.5)
    o.set(VEL, np.concatenate([ux, uy]))
    o.enforce_boundary()
    nx, ny = face_split(g, o.get(VEL))
    # x-face (1, 8) sits in the left solid band, normal (+1, 0) into the fluid
    assert nx[1 + 17 * 8] == pytest.approx(0.0, abs=1e-15)
    assert ny[1 + 16 * 8] == pytest.approx(0.1, abs=1e-15)      # 0.5 - 0.8 * 0.5
    # interior faces are untouched
    assert nx[8 + 17 * 8] == -1.0 and ny[8 + 16 * 8] == 0.5


# ----------------------------------------------------- restart check ----
def test_check_refmap_detects_shear():
    g, o = mk(24, 2)
    scenes.dam_break(o, g)
    o.initialize(tol=1e-10)
    ok, tot, cnt = o.check_refmap()
    assert ok and tot > 0 and cnt == 0
    J = np.tile(np.array([[1.0, 0.0], [3.0, 1.0]]).ravel(order="F"), 24 * 24)   # columns at ~18 degrees
    o.set(REFMAP_GRAD, J)
    ok, tot2, cnt2 = o.check_refmap()
    assert not ok and cnt2 == tot2 == tot


# ------------------------------------------------- computeFirst<4> ----
def quartic_field(x):
    """A polynomial of degree 4 in each coordinate and its gradient."""
    X, Y = x[:, 0], x[:, 1]
    Z = x[:, 2] if x.shape[1] == 3 else np.zeros_like(X)
    f = X ** 4 - 2 * X ** 3 * Y + 3 * Y ** 4 + X ** 2 * Z ** 2 - Z ** 3 + 0.5 * Y * Z - 0.25 * Z ** 4 * X + 1.0
    g = np.stack([4 * X ** 3 - 6 * X ** 2 * Y + 2 * X * Z ** 2 - 0.25 * Z ** 4,
                  -2 * X ** 3 + 12 * Y ** 3 + 0.5 * Z,
                  2 * X ** 2 * Z - 3 * Z ** 2 + 0.5 * Y - X * Z ** 3], axis=1)
    return f, g[:, :x.shape[1]]


@pytest.mark.parametrize("dim", [2, 3])
def test_first4_exact_for_quartics(dim):
    g, o = mk(14, dim)
    x = scenes.cell_positions(g)
    f, grad = quartic_field(x)
    o.set(PHI, f)
    o.restart()                     # phi_ref = phi, grad phi_ref = computeFirst<4>(phi_ref)
    got = o.get(PHI_REF_GRAD).reshape(-1, dim)
    m = interior_mask(g, 2)
    assert np.max(np.abs(got[m] - grad[m])) <= 1e-10 * np.max(np.abs(grad[m]))
    # next to a wall the clamped stencil is not exact any more
    assert np.max(np.abs(got[~m] - grad[~m])) > 1e-3


# values of (v0 - 8 v1 + 8 v3 - v4) / 12 with clamped indices, by hand, for
# v = i^2 on a 6-cell x axis and v = 10 j on a 5-cell y axis (unit spacing)
FIRST4_EDGE_X = [1 / 3, 23 / 12, 4.0, 6.0, 107 / 12, 14 / 3]
FIRST4_EDGE_Y = [5.0, 130 / 12, 10.0, 130 / 12, 5.0]


def test_first4_clamped_edges_by_hand():
    o = O.Oracle(2, [6, 5], 1.0, [0.0, 0.0], 0)
    i, j = np.meshgrid(np.arange(6), np.arange(5))      # x fastest after ravel
    o.set(PHI, (i ** 2 + 10.0 * j).ravel())
    o.restart()
    got = o.get(PHI_REF_GRAD).reshape(5, 6, 2)
    for jj in range(5):
        assert np.allclose(got[jj, :, 0], FIRST4_EDGE_X, rtol=0, atol=1e-13)
    for ii in range(6):
        assert np.allclose(got[:, ii, 1], FIRST4_EDGE_Y, rtol=0, atol=1e-13)


# ------------------------------------------- solid inside the domain ----
def _square_solid_fraction(c0, c1, d):
    """Fraction of the unit square {(s, t)} with c0 s + c1 t < d, exact."""
    from test_volume_oracle import half_plane_area
    return half_plane_area([c0, c1], d)


def test_solid_plane_face_fractions_closed_form():
    """Boundary::unions(ImplicitPlane) + finish: the node distances are
    linear, so getFaceFraction's four triangles give the exact solid area of
    each face (set to 1 above 0.9), and every face normal is the plane's."""
    g, o = mk(12, 3)
    p, n = np.array([0.0, 0.3, 0.0]), np.array([-0.3, 1.0, 0.2])
    o.solid_plane(p, n)
    nn = n / np.linalg.norm(n)
    fr, nr = face_split(g, o.get(FRAC)), face_split(g, o.get(NORMAL))
    dx, bw = g["dx"], g["bw"]
    partial = 0
    for a in range(3):
        pos = scenes.face_positions(g, a)
        cnt = [r + (1 if b == a else 0) for b, r in enumerate(g["res"])]
        idx = np.indices(cnt[::-1]).reshape(3, -1)[::-1]
        wall = (idx[a] <= bw) | (idx[a] >= g["res"][a] - bw)
        for b in range(3):
            wall |= (idx[b] < bw) | (idx[b] >= cnt[b] - bw)
        u, v = (a + 1) % 3, (a + 2) % 3
        for k in np.nonzero(~wall)[0]:
            # face square: corner at pos - dx/2 along u and v
            lo = pos[k].copy()
            lo[u] -= dx / 2
            lo[v] -= dx / 2
            nodes = [lo + s * dx * np.eye(3)[u] + t * dx * np.eye(3)[v] for s in (0, 1) for t in (0, 1)]
            if any(np.dot(q - p, nn) > np.min(np.minimum(q, 1 - q)) for q in nodes):
                continue   # a wall is nearer than the plane at a node: the distance is not linear there
            d = float(np.dot(p - lo, nn)) / dx       # solid: nn_u s + nn_v t < d (s, t in [0, 1])
            want = _square_solid_fraction(nn[u], nn[v], d)
            want = 1.0 if want > 0.9 else want
            assert fr[a][k] == pytest.approx(want, abs=1e-12), (a, k)
            partial += 0 < want < 1
            if np.dot(pos[k] - p, nn) < np.min(np.minimum(pos[k], 1 - pos[k])):
                assert nr[a][k] == pytest.approx(nn[a], abs=1e-15)
    assert partial > 60


def test_projection_partial_fractions_and_solid_velocity_by_hand():
    """One liquid cell with faces of fraction 0.25, 0, 0.6 and 1 (-x, +x, -y,
    +y), air around it at theta = 1/2, and a moving solid: Projector.cpp:30-68
    by hand.  rhs = -0.75*1 - 0.25*0.4 + 1*2 - 0.4*(-0.5) - 0.6*(-2) + 1*0.8
    = 3.35, diag = (0.75 + 1 + 0.4) * 2 = 4.3."""
    o = O.Oracle(2, [5, 5], 1.0, [0.0, 0.0], 1)
    p = hand_projection(o, "oracle")
    assert p == pytest.approx(3.35 / 4.3, rel=1e-14)


def hand_projection(sim, kind):
    """Sets the single-cell scene of the hand projection on `sim` (oracle or
    GPU), projects, checks the faces and returns the cell's pressure."""
    phi = np.full(25, 0.5)
    cid = 2 + 5 * 2
    phi[cid] = -0.5
    fx, fy = np.zeros(30), np.zeros(30)          # x faces 6 x 5, y faces 5 x 6
    ux, uy = np.zeros(30), np.zeros(30)
    sx, sy = np.zeros(30), np.zeros(30)
    fx[14], ux[14], sx[14] = 0.25, 1.0, 0.4      # -x face of cell (2, 2)
    fx[15], ux[15], sx[15] = 0.0, 2.0, 7.0       # +x (solid velocity unused: weight 1)
    fy[12], uy[12], sy[12] = 0.6, -0.5, -2.0     # -y
    fy[17], uy[17], sy[17] = 1.0, 3.0, 0.8       # +y (a wall: weight 0)
    sim.set(PHI, phi)
    sim.set(FRAC, np.concatenate([fx, fy]))
    sim.set(BVEL, np.concatenate([sx, sy]))
    sim.set(VEL, np.concatenate([ux, uy]))
    if kind == "oracle":
        r = sim.project(precond=O.PRECOND_JACOBI, tol=1e-14)
    else:
        r = sim.project(tol=1e-14)
    assert r["unknowns"] == 1
    p = sim.get(PRES)[cid]
    u = sim.get(VEL)
    nux, nuy = u[:30], u[30:]
    # Projector.cpp:91-123: one unknown side, theta = 1/2 on every face
    assert nux[15] == pytest.approx(2.0 - 2 * p, abs=1e-14)
    assert nux[14] == pytest.approx(1.0 + 2 * p, abs=1e-14)
    assert nuy[12] == pytest.approx(-0.5 + 2 * p, abs=1e-14)
    assert nuy[17] == pytest.approx(3.0 - 2 * p, abs=1e-14)
    return p


def hand_enforce(sim):
    """Boundary::enforce with uniform normal (0.6, 0.8), solid velocity
    (0.5, 0.25) and fluid velocity (1, -2): vn = -1.5 < 0, vt = (1.4, -1.05),
    so a face with 0.1 < fraction <= 1 becomes vel - vn n - 0.8 vt =
    (0.78, 0.04) in its axis; fractions at or below 0.1 are left alone."""
    fr = sim.get(FRAC).copy()
    nxf = 17 * 16
    fr[8 + 17 * 8] = 0.5          # interior x face, partial
    fr[9 + 17 * 8] = 0.05         # interior x face, below 0.1
    fr[nxf + 8 + 16 * 8] = 0.35   # interior y face, partial
    sim.set(FRAC, fr)
    sim.set(NORMAL, np.concatenate([np.full(nxf, 0.6), np.full(16 * 17, 0.8)]))
    sim.set(BVEL, np.concatenate([np.full(nxf, 0.5), np.full(16 * 17, 0.25)]))
    sim.set(VEL, np.concatenate([np.full(nxf, 1.0), np.full(16 * 17, -2.0)]))
    sim.enforce_boundary()
    u = sim.get(VEL)
    hit = fr > 0.1
    want = np.where(hit, np.concatenate([np.full(nxf, 0.78), np.full(16 * 17, 0.04)]),
                    np.concatenate([np.full(nxf, 1.0), np.full(16 * 17, -2.0)]))
    assert np.max(np.abs(u - want)) <= 1e-14
    assert hit[8 + 17 * 8] and hit[nxf + 8 + 16 * 8] and not hit[9 + 17 * 8]


def test_enforce_partial_fraction_branch_by_hand():
    g, o = mk(16, 2)
    hand_enforce(o)


# ------------------------------------------------- Hybrid-Newton restart ----
def _newton_plane(dim, n, scale):
    g = scenes.grid(n, dim)
    o = O.Oracle(dim, g["res"], g["dx"], g["origin"], g["bw"])
    x = scenes.cell_positions(g)
    nrm = np.array([0.6, 0.8] if dim == 2 else [0.48, 0.64, 0.6])
    sd = x @ nrm - 0.45
    o.set(PHI, sd)                                        # the cubic view's guesses
    o.set(PHI_ORI, scale * sd)                            # what Newton and the march work on
    o.set(PHI_GRAD, np.tile(scale * nrm, (len(sd), 1)).ravel())
    o.restart_newton()
    return g, o, sd, nrm


@pytest.mark.parametrize("dim", [2, 3])
def test_newton_on_a_plane_returns_the_exact_closest_point(dim):
    """Cubic Hermite reproduces a linear level set, so the modified Newton
    iteration lands on the exact closest point: the interface band (cells
    next to a sign change and one layer more) gets the exact signed
    distance and the unit normal even from a level set scaled by 2."""
    g, o, sd, nrm = _newton_plane(dim, 40 if dim == 2 else 24, 2.0)
    pr, gr = o.get(PHI_REF), o.get(PHI_REF_GRAD).reshape(-1, dim)
    # within 0.6 dx of the plane every cell has a neighbour across it (the
    # largest normal component is 0.64 or 0.8): all Newton cells
    band = (np.abs(sd) <= 0.6 * g["dx"]) & interior_mask(g, 4)
    assert band.sum() > 10
    assert np.max(np.abs(pr[band] - sd[band])) <= 1e-13
    assert np.max(np.abs(gr[band] - nrm)) <= 1e-13
    # the characteristic march carries the distance outward: two cells on
    # the band is exact too
    near = (np.abs(sd) <= 3.0 * g["dx"]) & interior_mask(g, 6)
    assert np.max(np.abs(pr[near] - sd[near])) <= 1e-6 * g["dx"]
    # the reference map is reset
    assert np.allclose(o.get(REFMAP).reshape(-1, dim), scenes.cell_positions(g), atol=0, rtol=0)


def test_newton_march_keeps_a_distance_field():
    """A signed distance (|grad phi| = 1) is a fixed point of the
    characteristic steps phi(x) <- phi(x - s dt n) + s dt, away from the walls
    whose clamped data bends the interpolant."""
    g, o, sd, nrm = _newton_plane(2, 64, 1.0)
    pr = o.get(PHI_REF)
    m = interior_mask(g, 24)
    assert np.max(np.abs(pr[m] - sd[m])) <= 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_newton_on_a_sphere_matches_the_distance(dim):
    """Exact sphere distance and normal as phi_ori and grad phi: the Newton
    band matches the true distance to the Hermite interpolant's accuracy."""
    n = 48 if dim == 2 else 32
    g = scenes.grid(n, dim)
    o = O.Oracle(dim, g["res"], g["dx"], g["origin"], g["bw"])
    x = scenes.cell_positions(g)
    c = np.full(dim, 0.5)
    r = np.linalg.norm(x - c, axis=1)
    sd = r - 0.3
    o.set(PHI, sd)
    o.set(PHI_ORI, sd)
    o.set(PHI_GRAD, ((x - c) / r[:, None]).ravel())
    o.restart_newton()
    pr, gr = o.get(PHI_REF), o.get(PHI_REF_GRAD).reshape(-1, dim)
    band = np.abs(sd) <= 1.5 * g["dx"]
    assert np.max(np.abs(pr[band] - sd[band])) <= 1e-3 * g["dx"]
    assert np.max(np.abs(gr[band] - (x[band] - c) / r[band, None])) <= 1e-3
    # a plain restart keeps phi as it is; Newton moves no band value by more
    # than the interpolant's error
    assert np.max(np.abs(pr[band] - sd[band])) < np.max(np.abs(sd[band])) * 1e-2


def test_newton_marks_sign_changes_and_one_layer():
    """prepareInterface + extrapolate(1) (HybridNewton.cpp:64-79, 103-123):
    with phi_ori = x - 0.5 on a 10-cell row the cells next to the sign change
    and their neighbours are Newton cells; the rest are marched."""
    o = O.Oracle(2, [10, 6], 0.1, [0.0, 0.0], 0)
    i, j = np.meshgrid(np.arange(10), np.arange(6))
    xpos = (i + 0.5) * 0.1
    phi = (xpos - 0.52).ravel()            # sign change between i = 4 and i = 5
    o.set(PHI, phi)
    o.set(PHI_ORI, phi)
    o.set(PHI_GRAD, np.tile([1.0, 0.0], phi.size))
    o.restart_newton()
    pr = o.get(PHI_REF).reshape(6, 10)
    # the band i = 3..6 is Newton-exact; the march reproduces the rest of a distance field as well
    assert np.allclose(pr, (xpos - 0.52), atol=1e-13, rtol=0)
