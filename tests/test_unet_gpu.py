"""The BF16 MFMA network on the GPU against the FP64 network oracle.

Two bars per quantity: against the oracle evaluated with the GPU's BF16
rounding points (oracle q=bf16; the remaining differences are FP32
accumulation order and the rare BF16 rounding that flips with it), and a
looser one against the exact FP64 network, which bounds the BF16 error
itself.  Grid sizes are deliberately unequal and not multiples of the
convolution tiles (16 x 4 x 4 voxels for 32 and more output channels,
16 x 8 x 8 for 16) so that partial tiles and axis order are tested; the
larger shape holds whole 16 x 8 x 8 tiles beside partial ones on the finest
level.
"""
import numpy as np
import pytest

import unet_oracle as UO
from garm_amd import capi
from helpers import say

pytestmark = pytest.mark.gpu

SHAPE = (12, 8, 20)   # (nz, ny, nx)
LARGE = (24, 20, 36)


def _rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _case(seed, shape=SHAPE):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    blob = ((x - nx * 0.45) / (nx * 0.4)) ** 2 + ((y - ny * 0.5) / (ny * 0.45)) ** 2 + ((z - nz * 0.4) / (nz * 0.5)) ** 2
    mask = (blob < 1.0).astype(np.uint8)
    r = (rng.normal(size=shape) * mask).astype(np.float32)
    return rng, r, mask


def _params(seed):
    P = UO.he_init(seed)
    rng = np.random.default_rng(seed + 100)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.normal(0.0, 0.05, size=P[k].shape)
    # the GPU holds FP32 master weights: compare against the same values
    return UO.unflatten(UO.flatten(P).astype(np.float32))


def test_param_layout_round_trip():
    say("unet param round trip")
    net = capi.Unet(seed=1)
    assert capi.Unet.param_count() == UO.flatten(UO.he_init(0)).size
    p0 = net.get_params()
    assert np.all(np.isfinite(p0)) and np.any(p0 != 0)
    P = _params(4)
    flat = UO.flatten(P).astype(np.float32)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)


@pytest.mark.parametrize("seed,shape", [(11, SHAPE), (12, SHAPE), (13, LARGE)])
def test_forward_matches_oracle(seed, shape):
    say(f"unet forward seed={seed} shape={shape}")
    _, r, mask = _case(seed, shape)
    P = _params(seed)
    net = capi.Unet(seed=0)
    net.set_params(UO.flatten(P))
    y = net.forward(r, mask)
    x = UO.make_input(r.astype(np.float64), mask.astype(np.float64))
    y_q = UO.forward(P, x, mask, q=UO.bf16)
    y_e = UO.forward(P, x, mask)
    e_q, e_e = _rel_l2(y, y_q), _rel_l2(y, y_e)
    say(f"  rel L2 vs bf16-rounded oracle {e_q:.2e}, vs fp64 oracle {e_e:.2e}")
    assert np.all(y[mask == 0] == 0)
    assert e_q <= 1.5e-2
    assert e_e <= 4e-2


@pytest.mark.parametrize("shape", [SHAPE, LARGE])
def test_loss_gradient_matches_oracle(shape):
    say(f"unet loss gradient shape={shape}")
    rng, r, mask = _case(21, shape)
    target = (rng.normal(size=shape) * mask).astype(np.float32)
    P = _params(21)
    net = capi.Unet(seed=0)
    net.set_params(UO.flatten(P))
    loss, g = net.loss_grad(r, mask, target)
    x = UO.make_input(r.astype(np.float64), mask.astype(np.float64))
    loss_q, g_q = UO.loss_and_grad(P, x, mask.astype(np.float64), target.astype(np.float64), q=UO.bf16)
    loss_e, g_e = UO.loss_and_grad(P, x, mask.astype(np.float64), target.astype(np.float64))
    assert abs(loss - loss_q) <= 2e-3 * loss_q
    assert abs(loss - loss_e) <= 2e-2 * loss_e
    G = UO.unflatten(g)
    worst_q = worst_e = 0.0
    for name in G:
        if name == "L1a.w":   # input channels 2, 3 are always zero
            assert np.all(G[name][:, 2:, :] == 0)
        e_q, e_e = _rel_l2(G[name], g_q[name]), _rel_l2(G[name], g_e[name])
        say(f"  {name:6s} rel L2 vs bf16 oracle {e_q:.2e}, vs fp64 {e_e:.2e}")
        worst_q, worst_e = max(worst_q, e_q), max(worst_e, e_e)
    assert worst_q <= 2e-2
    assert worst_e <= 1e-1


def test_training_reduces_the_loss_on_a_fixed_pair():
    say("unet training on one pair")
    _, r, mask = _case(31)
    # a target that is linear in the input (training flips the pair's sign):
    # the 3x3x3 box average of r, which the network can represent
    rp = np.pad(r.astype(np.float64), 1)
    nz, ny, nx = SHAPE
    p = sum(rp[dz:dz + nz, dy:dy + ny, dx:dx + nx] for dz in range(3) for dy in range(3) for dx in range(3)) / 27.0
    p = (p * mask).astype(np.float32)
    net = capi.Unet(seed=5)
    hist = net.train(r[None], p[None], mask[None], steps=300, lr=2e-3, seed=1)
    say(f"  loss {hist[:10].mean():.3e} -> {hist[-10:].mean():.3e}")
    assert np.all(np.isfinite(hist))
    assert hist[-10:].mean() < 0.2 * hist[:10].mean()


def test_bad_arguments_raise():
    net = capi.Unet(seed=0)
    with pytest.raises(capi.GarmError):
        net.forward(np.zeros((6, 8, 8), np.float32), np.ones((6, 8, 8), np.uint8))   # 6 is not a multiple of 4
    with pytest.raises(capi.GarmError):
        net.set_params(np.zeros(5, np.float32))
