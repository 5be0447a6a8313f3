"""Pins of the FP64 network oracle (oracle/unet_oracle.py) on the CPU.

The network has no counterpart in the reference, so its oracle is pinned by
its own definition: the convolution against a brute-force loop, pooling and
upsampling against their adjoint identities, the hand-written backward pass
against central finite differences of the forward pass, and the BF16
rounding against known answers.
"""
import numpy as np

import unet_oracle as U


def test_conv3_matches_brute_force():
    rng = np.random.default_rng(1)
    nz, ny, nx, ci, co = 3, 4, 5, 3, 2
    x = rng.normal(size=(nz, ny, nx, ci))
    w = rng.normal(size=(27, ci, co))
    b = rng.normal(size=co)
    y = U.conv3(x, w, b)
    ref = np.zeros((nz, ny, nx, co)) + b
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            sz, sy, sx = z + dz, yy + dy, xx + dx
                            if 0 <= sz < nz and 0 <= sy < ny and 0 <= sx < nx:
                                t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                ref[z, yy, xx] += x[sz, sy, sx] @ w[t]
    assert np.allclose(y, ref, rtol=0, atol=1e-12)


def test_pool_and_up_are_adjoint_up_to_eight():
    # <pool(a), b> = <a, up(b)> / 8 and pool(up(b)) = b
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 6, 8, 3))
    b = rng.normal(size=(2, 3, 4, 3))
    assert np.isclose(np.sum(U.pool(a) * b), np.sum(a * U.up(b)) / 8.0)
    assert np.allclose(U.pool(U.up(b)), b)
    assert np.allclose(U.pool_backward(b), U.up(b) / 8.0)
    assert np.allclose(U.up_backward(U.up(b)), 8.0 * b)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = U.he_init(5)
    for k in P:   # nonzero biases so that every bias gradient is exercised
        if k.endswith(".b"):
            P[k] = rng.normal(0.0, 0.1, size=P[k].shape)
    shape = (4, 4, 8)
    r = rng.normal(size=shape)
    mask = (rng.random(shape) < 0.7).astype(np.float64)
    x = U.make_input(r, mask)
    target = rng.normal(size=shape)
    _, g = U.loss_and_grad(P, x, mask, target)
    h = 1e-6
    checked = 0
    for name in ["L1a.w", "L1b.b", "L2a.w", "L3b.w", "L3a.b", "L4a.w", "L5a.w", "L5b.b", "out.w", "out.b"]:
        flat = P[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(3, flat.size), replace=False)
        if name == "L1a.w":   # only live input channels carry a gradient
            idx = [t * 4 * 16 + c * 16 + o for t, c, o in ((0, 0, 3), (13, 1, 7), (26, 0, 15))]
        for i in idx:
            old = flat[i]
            flat[i] = old + h
            lp, _ = U.loss_and_grad(P, x, mask, target)
            flat[i] = old - h
            lm, _ = U.loss_and_grad(P, x, mask, target)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, an)
            checked += 1
    assert checked >= 25


def test_bf16_rounding_known_answers():
    e = 2.0 ** -8
    vals = np.array([1.0, 1.0 + e, 1.0 + 3 * e, -2.5, 1.0 + e + 2.0 ** -20, 3.0e38 * 0 + 65504.0])
    want = np.array([1.0, 1.0, 1.0 + 4 * e, -2.5, 1.0 + 2 * e, 65536.0])
    assert np.array_equal(U.bf16(vals), want)


def test_flatten_round_trip():
    P = U.he_init(9)
    flat = U.flatten(P)
    n = sum(27 * ci * co + co for _, ci, co in U.LAYERS) + 17
    assert flat.size == n
    Q = U.unflatten(flat)
    assert all(np.array_equal(P[k], Q[k]) for k in P)
