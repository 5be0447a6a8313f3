"""FP64 numpy restatement of the learned pressure preconditioner network.

ORACLE — test infrastructure only (tests/ and bench.py's checker).  It has
no reference counterpart: the reference solves with Eigen's IC-PCG
(LinearSystem.h:10-39); the network is the preconditioner the GPU path adds.

Architecture (channels-last tensors of shape (nz, ny, nx, C)):
  input  (r, mask, 0, 0)                                 4 ch, level 1
  L1a  conv3 4->16  + ReLU      L1b conv3 16->16 + ReLU  = e1
  pool (2x2x2 mean)  ->  L2a conv3 16->32 + ReLU, L2b conv3 32->32 + ReLU = e2
  pool               ->  L3a conv3 32->64 + ReLU, L3b conv3 64->64 + ReLU = b3
  up (nearest x2), concat [up(b3), e2] (96) -> L4a 96->32 + ReLU, L4b 32->32 + ReLU = d2
  up, concat [up(d2), e1] (48)             -> L5a 48->16 + ReLU, L5b 16->16 + ReLU = d1
  out  1x1 conv 16->1 (+bias), times mask
conv3 = 3x3x3 cross-correlation with zero padding; weights W[t][ci][co],
t = (dz+1)*9 + (dy+1)*3 + (dx+1).

forward() and backward() take an optional rounding function `q` applied
where the GPU network stores a BF16 tensor (bf16() below): the input, the
convolution weights, every activation, pooled and gradient tensor.  With the
default identity they are the exact FP64 network.
"""
import numpy as np

LAYERS = [  # name, cin, cout
    ("L1a", 4, 16), ("L1b", 16, 16),
    ("L2a", 16, 32), ("L2b", 32, 32),
    ("L3a", 32, 64), ("L3b", 64, 64),
    ("L4a", 96, 32), ("L4b", 32, 32),
    ("L5a", 48, 16), ("L5b", 16, 16),
]
IN_CHANNELS = 2   # live input channels (residual, mask); the rest are zero


def ident(x):
    return x


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def he_init(seed):
    """He-normal weights (std sqrt(2 / fan_in)), zero biases, seeded."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout in LAYERS:
        fan = 27 * (IN_CHANNELS if name == "L1a" else cin)
        w = rng.normal(0.0, np.sqrt(2.0 / fan), size=(27, cin, cout))
        if name == "L1a":
            w[:, IN_CHANNELS:, :] = 0.0
        params[name + ".w"] = w
        params[name + ".b"] = np.zeros(cout)
    params["out.w"] = rng.normal(0.0, np.sqrt(2.0 / 16), size=(16,))
    params["out.b"] = np.zeros(1)
    return params


def conv3(x, w, b):
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nz, ny, nx, w.shape[2])) + b
    t = 0
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                y += xp[dz:dz + nz, dy:dy + ny, dx:dx + nx, :] @ w[t]
                t += 1
    return y


def conv3_backward(x, w, gy):
    """Gradients of conv3 with respect to x, w and b."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    t = 0
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                sl = xp[dz:dz + nz, dy:dy + ny, dx:dx + nx, :]
                gw[t] = sl.reshape(-1, sl.shape[-1]).T @ gy.reshape(-1, gy.shape[-1])
                gxp[dz:dz + nz, dy:dy + ny, dx:dx + nx, :] += gy @ w[t].T
                t += 1
    return gxp[1:-1, 1:-1, 1:-1, :], gw, gy.sum(axis=(0, 1, 2))


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_backward(gy):
    return np.repeat(np.repeat(np.repeat(gy, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_backward(gy):
    nz, ny, nx, c = gy.shape
    return gy.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def make_input(r, mask):
    x = np.zeros(r.shape + (4,))
    x[..., 0] = r
    x[..., 1] = mask
    return x


def forward(params, x, mask, keep=False, q=ident):
    """x: (nz, ny, nx, 4); returns (nz, ny, nx) output and the tape."""
    P = params
    tape = {}

    def cr(name, inp):
        z = conv3(inp, q(P[name + ".w"]), P[name + ".b"])
        a = q(np.maximum(z, 0.0))
        tape[name] = (inp, z)
        return a

    a = cr("L1a", q(x))
    e1 = cr("L1b", a)
    a = cr("L2a", q(pool(e1)))
    e2 = cr("L2b", a)
    a = cr("L3a", q(pool(e2)))
    b3 = cr("L3b", a)
    c2 = np.concatenate([up(b3), e2], axis=-1)
    a = cr("L4a", c2)
    d2 = cr("L4b", a)
    c1 = np.concatenate([up(d2), e1], axis=-1)
    a = cr("L5a", c1)
    d1 = cr("L5b", a)
    y = (d1 @ P["out.w"] + P["out.b"][0]) * mask
    tape["out"] = d1
    return (y, tape) if keep else y


def backward(params, tape, mask, gy, q=ident):
    """Gradient of sum(gy * y) with respect to every parameter."""
    P = params
    g = {}
    d1 = tape["out"]
    gyo = gy * mask
    g["out.w"] = (d1 * gyo[..., None]).sum(axis=(0, 1, 2))
    g["out.b"] = np.array([gyo.sum()])
    gd1 = q(gyo[..., None] * P["out.w"][None, None, None, :])

    def cr_b(name, ga):
        inp, z = tape[name]
        gz = ga * (z > 0)
        gx, gw, gb = conv3_backward(inp, q(P[name + ".w"]), gz)
        g[name + ".w"] = gw
        g[name + ".b"] = gb
        return q(gx)

    ga = cr_b("L5b", gd1)
    gc1 = cr_b("L5a", ga)
    gd2 = q(up_backward(gc1[..., :32]))
    ge1 = gc1[..., 32:]
    ga = cr_b("L4b", gd2)
    gc2 = cr_b("L4a", ga)
    gb3 = q(up_backward(gc2[..., :64]))
    ge2 = gc2[..., 64:]
    ga = cr_b("L3b", gb3)
    gp2 = cr_b("L3a", ga)
    ge2 = q(ge2 + q(pool_backward(gp2)))
    ga = cr_b("L2b", ge2)
    gp1 = cr_b("L2a", ga)
    ge1 = q(ge1 + q(pool_backward(gp1)))
    ga = cr_b("L1b", ge1)
    cr_b("L1a", ga)
    g["L1a.w"][:, IN_CHANNELS:, :] = 0.0
    return g


def loss_and_grad(params, x, mask, target, q=ident):
    """Mean squared error over the masked cells and its parameter gradient."""
    y, tape = forward(params, x, mask, keep=True, q=q)
    m = mask.sum()
    diff = (y - target) * mask
    loss = (diff ** 2).sum() / m
    g = backward(params, tape, mask, 2.0 * diff / m, q=q)
    return loss, g


def flatten(params):
    """Parameters in the order and layout of garm_unet_get_params."""
    parts = []
    for name, _, _ in LAYERS:
        parts += [params[name + ".w"].ravel(), params[name + ".b"].ravel()]
    parts += [params["out.w"].ravel(), params["out.b"].ravel()]
    return np.concatenate(parts)


def unflatten(flat):
    params, off = {}, 0
    for name, cin, cout in LAYERS:
        n = 27 * cin * cout
        params[name + ".w"] = np.asarray(flat[off:off + n], dtype=np.float64).reshape(27, cin, cout)
        params[name + ".b"] = np.asarray(flat[off + n:off + n + cout], dtype=np.float64)
        off += n + cout
    params["out.w"] = np.asarray(flat[off:off + 16], dtype=np.float64)
    params["out.b"] = np.asarray(flat[off + 16:off + 17], dtype=np.float64)
    return params
