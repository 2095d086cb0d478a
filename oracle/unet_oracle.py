"""CPU oracle of the learned preconditioner's network (test infrastructure).

FP64 numpy restatement of the 4-level U-Net that similie_amd/csrc/unet.hip
and unet_net.cpp run in BF16/FP32 on the GPU: forward pass, parameter
gradient of a linear functional of the output, and the direction loss with
its gradient.  Used only by the tests, which check the GPU kernels against
it; its own gradients are pinned by central finite differences in
tests/test_unet_oracle.py.

The network has no counterpart in the reference (it is new work), so there
is no reference file to cite; the layout follows the parameter vector
documented in include/sil_hip.h (sil_unet_*).
"""
import numpy as np

C = 32
N_IN = 9 * C * 2 + C
N_LAYER = 9 * C * C + C
N_OUT = 9 * C + 1
N_LAYERS = 12
N_PARAMS = N_IN + N_LAYERS * N_LAYER + N_OUT


def layer_slices(l):
    """(W, b) slices of 32->32 layer l = 1..12 in the parameter vector."""
    base = N_IN + (l - 1) * N_LAYER
    return slice(base, base + 9 * C * C), slice(base + 9 * C * C, base + N_LAYER)


OUT_BASE = N_IN + N_LAYERS * N_LAYER


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    assert p.shape == (N_PARAMS,)
    win = p[: 9 * C * 2].reshape(9, C, 2)
    bin_ = p[9 * C * 2: N_IN]
    layers = []
    for l in range(1, N_LAYERS + 1):
        ws, bs = layer_slices(l)
        layers.append((p[ws].reshape(9, C, C), p[bs]))
    wout = p[OUT_BASE: OUT_BASE + 9 * C].reshape(9, C)
    bout = p[OUT_BASE + 9 * C]
    return win, bin_, layers, wout, bout


def conv(x, w, b):
    """3x3 zero-padded convolution, x [B,H,W,Ci], w [9][Co][Ci] (tap = 3 dy + dx)."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, W, w.shape[1]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        y += xp[:, dy:dy + H, dx:dx + W, :] @ w[t].T
    return y + b


def conv_bwd(x, w, dz):
    """(dx, dw, db) of conv for the pre-activation gradient dz."""
    B, H, W, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for t in range(9):
        dy, dx = divmod(t, 3)
        dxp[:, dy:dy + H, dx:dx + W, :] += dz @ w[t]
        dw[t] = np.einsum("bhwo,bhwc->oc", dz, xp[:, dy:dy + H, dx:dx + W, :])
    return dxp[:, 1:-1, 1:-1, :], dw, dz.sum(axis=(0, 1, 2))


def pool(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])


def pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(dy):
    return dy[:, 0::2, 0::2] + dy[:, 1::2, 0::2] + dy[:, 0::2, 1::2] + dy[:, 1::2, 1::2]


def relu(v):
    return np.maximum(v, 0.0)


# (layer, input name, output name) in forward order; "pool:" / "up:" entries are the
# parameter-free steps
GRAPH = [
    ("in", "x", "a0"),
    (1, "a0", "e0"),
    ("pool", "e0", "p1"),
    (2, "p1", "a1"), (3, "a1", "e1"),
    ("pool", "e1", "p2"),
    (4, "p2", "a2"), (5, "a2", "e2"),
    ("pool", "e2", "p3"),
    (6, "p3", "a3"), (7, "a3", "b3"),
    ("up", ("b3", "e2"), "u2"),
    (8, "u2", "c2"), (9, "c2", "f2"),
    ("up", ("f2", "e1"), "u1"),
    (10, "u1", "c1"), (11, "c1", "f1"),
    ("up", ("f1", "e0"), "u0"),
    (12, "u0", "c0"),
    ("out", "c0", "out"),
]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def conv32(x, w, b):
    """conv() accumulated in FP32 (as the matrix cores do)."""
    return conv(np.float32(x), np.float32(w), np.float32(b)).astype(np.float32).astype(np.float64)


def split_bf16(a):
    """a (as FP32) as the two-term BF16 sum hi + lo, hi = bf16(a),
    lo = bf16(a - hi): a to 16 significant bits, |a - hi - lo| <= 2^-16 |a|
    (a - hi is exact in FP32).  The GPU carries the input layer's FP32
    operands in this form (similie_amd/csrc/unet.hip, split_bf16)."""
    a = f32(a)
    hi = bf16(a)
    return hi + bf16(a - hi)


def conv_in_split32(x, w, b):
    """The first layer as the GPU forms it: every input value and weight
    replaced by its split (split_bf16), and per output channel the chain
    a = b; a = fp32(x~_t[ci] * w~[t][c][ci] + a) over taps t = 0..8 and input
    channels ci = 0, 1, each product exact in FP64 (two 16-bit significands)
    with one rounding to FP32 per step.  The GPU sums the same four exact
    BF16 products per x*w on its BF16 matrix cores (bias first, two taps per
    instruction, taps in order), which differs from this chain only in FP32
    rounding: after the BF16 rounding of the activation the two agree on
    every element of the layer-by-layer test but a handful."""
    B, H, W, _ = x.shape
    x = split_bf16(x)
    w = split_bf16(w)
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    a = np.broadcast_to(f32(b), (B, H, W, w.shape[1])).copy()
    for t in range(9):
        dy, dx = divmod(t, 3)
        for ci in range(2):
            a = f32(xp[:, dy:dy + H, dx:dx + W, ci][..., None] * w[t][:, ci] + a)
    return a


def conv_bwd32(x, w, dz):
    dx, dw, db = conv_bwd(np.float32(x), np.float32(w), np.float32(dz))
    return tuple(a.astype(np.float32).astype(np.float64) for a in (dx, dw, db))


# The same FP32 convolutions with every sum taken in another order: input
# channels permuted (the matrix products sum them differently), taps in
# reverse, pixels mirrored for the weight gradients.  Mathematically equal to
# conv32 / conv_bwd32; the difference between the two is what the order of
# FP32 accumulation alone does to a network with BF16 storage, the yardstick
# for how far a GPU kernel that sums in its own order may land from the
# emulating oracle.
_PERM = np.random.default_rng(12345).permutation(C)


def _perm(n):
    return _PERM if n == C else np.arange(n)[::-1]


def conv32_reordered(x, w, b):
    x32, w32 = np.float32(x), np.float32(w)
    pi = _perm(x.shape[-1])
    x32, w32 = x32[..., pi], w32[:, :, pi]
    B, H, W, _ = x.shape
    xp = np.pad(x32, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, W, w.shape[1]), dtype=np.float32)
    for t in reversed(range(9)):
        dy, dx = divmod(t, 3)
        y += xp[:, dy:dy + H, dx:dx + W, :] @ w32[t].T
    return (y + np.float32(b)).astype(np.float64)


def conv_bwd32_reordered(x, w, dz):
    x32, w32, dz32 = np.float32(x), np.float32(w), np.float32(dz)
    po = _perm(dz.shape[-1])
    B, H, W, _ = x.shape
    xp = np.pad(x32, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w32)
    for t in reversed(range(9)):
        dy, dx = divmod(t, 3)
        dxp[:, dy:dy + H, dx:dx + W, :] += dz32[..., po] @ w32[t][po, :]
        xs = xp[:, dy:dy + H, dx:dx + W, :][:, ::-1, ::-1, :]
        dw[t] = np.einsum("bhwo,bhwc->oc", dz32[:, ::-1, ::-1, :], xs)
    db = dz32[:, ::-1, ::-1, :].sum(axis=(0, 1, 2), dtype=np.float32)
    return tuple(a.astype(np.float32).astype(np.float64) for a in (dxp[:, 1:-1, 1:-1, :], dw, db))


def forward(params, x, emulate=False, reordered=False, a0=None):
    """Network output [B,H,W] and the activations, x [B,H,W,2]; with a0 the
    input layer's activation is taken as given (a GPU's, say) and the rest
    computed from it.

    emulate=True computes as the GPU kernels do: FP32 input and first/last
    layer weights, BF16 weights of the 32->32 layers, BF16 activations and
    FP32 accumulation (the first layer on two-term BF16 splits of its FP32
    input and weights, summed in FP32 in the GPU's tap order).  BF16 storage makes a network of this depth sensitive
    to the last bit of each accumulation (one in ten activations of the
    coarse levels rounds to the other neighbouring BF16 value when only the
    accumulation order changes), so only an oracle that accumulates in FP32
    too can be compared tightly with the GPU."""
    q = bf16 if emulate else (lambda a: a)
    qw = f32 if emulate else (lambda a: a)
    cv = (conv32_reordered if reordered else conv32) if emulate else conv
    win, bin_, layers, wout, bout = unpack(params)
    act = {"x": qw(np.asarray(x, dtype=np.float64))}
    for op, src, dst in GRAPH:
        if op == "in" and a0 is not None:
            act[dst] = np.asarray(a0, dtype=np.float64)
        elif op == "in":
            act[dst] = q(relu(conv_in_split32(act[src], win, bin_) if emulate else conv(act[src], win, bin_)))
        elif op == "pool":
            act[dst] = q(qw(pool(act[src])))
        elif op == "up":
            act[dst] = q(up(act[src[0]]) + act[src[1]])
        elif op == "out":
            act[dst] = cv(act[src], qw(wout)[:, None, :], qw(np.array([bout])))[..., 0]
        else:
            w, b = layers[op - 1]
            act[dst] = q(relu(cv(act[src], q(w), qw(b))))
    return act["out"], act


def backward(params, act, dout, emulate=False, reordered=False):
    """Gradient of sum(dout * out) with respect to the parameter vector
    (emulate=True: BF16 storage of the activation gradients, as on the GPU;
    reordered=True: every FP32 sum in another order)."""
    q = bf16 if emulate else (lambda a: a)
    qw = f32 if emulate else (lambda a: a)
    cb = (conv_bwd32_reordered if reordered else conv_bwd32) if emulate else conv_bwd
    win, bin_, layers, wout, bout = unpack(params)
    g = np.zeros(N_PARAMS)
    grad = {k: np.zeros_like(v) for k, v in act.items() if k != "out"}
    for op, src, dst in reversed(GRAPH):
        if op == "out":
            dx, dw, db = cb(act[src], qw(wout)[:, None, :], qw(np.asarray(dout, dtype=np.float64))[..., None])
            g[OUT_BASE: OUT_BASE + 9 * C] = dw[:, 0, :].ravel()
            g[OUT_BASE + 9 * C] = db[0]
            grad[src] = q(grad[src] + dx)
        elif op == "up":
            grad[src[0]] = q(qw(grad[src[0]] + up_bwd(grad[dst])))
            grad[src[1]] = q(grad[src[1]] + grad[dst])
        elif op == "pool":
            grad[src] = q(qw(grad[src] + pool_bwd(grad[dst])))
        else:
            dz = grad[dst] * (act[dst] > 0)
            if op == "in":
                dx, dw, db = cb(act[src], qw(win), dz)
                g[: 9 * C * 2] = dw.ravel()
                g[9 * C * 2: N_IN] = db
            else:
                w, _ = layers[op - 1]
                dx, dw, db = cb(act[src], q(w), dz)
                ws, bs = layer_slices(op)
                g[ws] = dw.ravel()
                g[bs] = db
            grad[src] = q(grad[src] + dx)
    return g


def apply5(u, wx, wy):
    """Dirichlet-zero 5-point operator with edge weights (interior rows only)."""
    u = np.asarray(u, dtype=np.float64)
    B, n, _ = u.shape
    out = np.zeros_like(u)
    v = u.copy()
    v[:, 0, :] = v[:, -1, :] = 0.0
    v[:, :, 0] = v[:, :, -1] = 0.0
    c = v[:, 1:-1, 1:-1]
    out[:, 1:-1, 1:-1] = (wx[:, 1:-1, 1:-1] * (c - v[:, 2:, 1:-1]) + wx[:, :-2, 1:-1] * (c - v[:, :-2, 1:-1])
                          + wy[:, 1:-1, 1:-1] * (c - v[:, 1:-1, 2:]) + wy[:, 1:-1, :-2] * (c - v[:, 1:-1, :-2]))
    return out


def direction_loss(N, r, e, wx, wy):
    """Per-image loss 1 - (r.N)^2/((e.r)(N.A N)) and the gradient of its batch mean."""
    N = np.asarray(N, dtype=np.float64)
    AN = apply5(N, wx, wy)
    B = N.shape[0]
    a = np.einsum("bij,bij->b", r, N)
    b = np.einsum("bij,bij->b", N, AN)
    E = np.einsum("bij,bij->b", e, r)
    loss = 1.0 - a * a / (E * b)
    dN = (2 * a / (E * b * B))[:, None, None] * (-r + (a / b)[:, None, None] * AN)
    return loss, dN
