"""FP64 numpy restatement of the pressure-correction network (test
infrastructure: the checker of the HIP forward and training kernels in
k_learned.hip; only tests/ load it).

Layout: activations are (nx, ny, nz, C) arrays (channels last, as on the
GPU), convolutions are 3x3x3 'same' with zero padding, weights are
W[co, tap, ci] with tap = (dx+1) + 3 (dy+1) + 9 (dz+1).

Architecture (4 levels, 16/32/48/64 channels; the 3 input fields are
average-pooled by 4 per axis first, so level 1 is N/4):
    x0 = pool4(input: log k, p, r) padded to 8 channels             N/4
    s1 = relu(conv(x0, 8->16))                                     N/4
    s2 = relu(conv(pool(s1), 16->32))                              N/8
    s3 = relu(conv(pool(s2), 32->48))                              N/16
    b4 = relu(conv(pool(s3), 48->64))                              N/32
    d3 = relu(conv(cat(up(b4), s3), 112->48))                      N/16
    d2 = relu(conv(cat(up(d3), s2), 80->32))                       N/8
    d1 = relu(conv(cat(up(d2), s1), 48->16))                       N/4
    y  = conv(d1, 16->1)                                           N/4
Loss (scale invariant) against the target pooled by 4:
    L = 1 - (y.t)^2 / ((y.y)(t.t)).
"""
from __future__ import annotations

import numpy as np

LAYERS = [("e1", 8, 16), ("e2", 16, 32), ("e3", 32, 48), ("e4", 48, 64),
          ("d3", 112, 48), ("d2", 80, 32), ("d1", 48, 16), ("out", 16, 1)]
CIN0 = 8


def bf16(a):
    """Round to the nearest BF16 value, ties to even (float64 result): the
    rounding the GPU applies where it stores weights, activations and the
    backward pass's incoming gradients in BF16."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    P = {}
    for name, ci, co in LAYERS:
        std = np.sqrt(2.0 / (27 * ci))
        P[name + ".w"] = rng.normal(0.0, std, (co, 27, ci))
        P[name + ".b"] = np.zeros(co)
    return P


def _pad(x):
    return np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))


def conv(x, W, b):
    nx, ny, nz, ci = x.shape
    co = W.shape[0]
    xp = _pad(x)
    out = np.broadcast_to(b, (nx, ny, nz, co)).copy()
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                tap = dx + 3 * dy + 9 * dz
                sl = xp[dx:dx + nx, dy:dy + ny, dz:dz + nz, :]
                out += sl @ W[:, tap, :].T
    return out


def conv_backward(x, W, gout):
    """Gradients of conv(x, W, b) given dL/dout: returns (dx, dW, db)."""
    nx, ny, nz, ci = x.shape
    xp = _pad(x)
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                tap = dx + 3 * dy + 9 * dz
                sl = xp[dx:dx + nx, dy:dy + ny, dz:dz + nz, :]
                dW[:, tap, :] = np.einsum("xyzo,xyzi->oi", gout, sl)
                dxp[dx:dx + nx, dy:dy + ny, dz:dz + nz, :] += gout @ W[:, tap, :]
    return dxp[1:-1, 1:-1, 1:-1, :], dW, gout.sum(axis=(0, 1, 2))


def pool(x):
    nx, ny, nz, c = x.shape
    return x.reshape(nx // 2, 2, ny // 2, 2, nz // 2, 2, c).mean(axis=(1, 3, 5))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_backward(g):
    nx, ny, nz, c = g.shape
    return g.reshape(nx // 2, 2, ny // 2, 2, nz // 2, 2, c).sum(axis=(1, 3, 5))


def make_input(fields):
    """fields: (3, nx, ny, nz) planar -> pooled-by-4 (nx/4, ny/4, nz/4, 8)."""
    x = np.moveaxis(np.asarray(fields, dtype=np.float64), 0, -1)
    x = pool(pool(x))
    out = np.zeros(x.shape[:3] + (CIN0,))
    out[..., :3] = x
    return out


def pool4_target(t):
    return pool(pool(np.asarray(t, dtype=np.float64)[..., None]))


def forward(P, x0, rounded=False):
    """x0: pooled 8-channel input (make_input).  Returns (y, cache).  With
    rounded=True the weights, the input and every stored activation (each
    layer's output and each pooled tensor) are rounded to BF16 as the GPU
    stores them; sums stay exact."""
    rnd = bf16 if rounded else (lambda a: a)
    C = {"rounded": rounded}
    x0 = rnd(x0)
    C["x0"] = x0
    def layer(name, x):
        z = conv(x, rnd(P[name + ".w"]), P[name + ".b"])
        C[name + ".in"] = x
        C[name + ".z"] = z
        return rnd(np.maximum(z, 0.0))
    s1 = layer("e1", x0)
    s2 = layer("e2", rnd(pool(s1)))
    s3 = layer("e3", rnd(pool(s2)))
    b4 = layer("e4", rnd(pool(s3)))
    d3 = layer("d3", np.concatenate([up(b4), s3], axis=-1))
    d2 = layer("d2", np.concatenate([up(d3), s2], axis=-1))
    d1 = layer("d1", np.concatenate([up(d2), s1], axis=-1))
    C["out.in"] = d1
    y = conv(d1, rnd(P["out.w"]), P["out.b"])
    return y, C


def loss(y, t):
    yt, yy, tt = float((y * t).sum()), float((y * y).sum()), float((t * t).sum())
    L = 1.0 - yt * yt / (yy * tt)
    gy = -2.0 * yt / (yy * tt) * t + 2.0 * yt * yt / (yy * yy * tt) * y
    return L, gy


def backward(P, C, gy):
    """Gradients of every weight and bias.  After a rounded forward, each
    layer's incoming gradient gz and its weights enter the convolution's
    backward in BF16 (the bias gradient sums the unrounded gz), as on the
    GPU."""
    G = {}
    rnd = bf16 if C.get("rounded") else (lambda a: a)

    def conv_back(name, gz):
        gx, G[name + ".w"], _ = conv_backward(C[name + ".in"], rnd(P[name + ".w"]), rnd(gz))
        G[name + ".b"] = gz.sum(axis=(0, 1, 2))
        return gx

    g = conv_back("out", gy)

    def layer_back(name, g_act):
        return conv_back(name, g_act * (C[name + ".z"] > 0))

    g = layer_back("d1", g)                   # g: d/d cat(up(d2), s1)
    g_up2, g_s1 = g[..., :32], g[..., 32:]
    g = layer_back("d2", up_backward(g_up2))
    g_up3, g_s2 = g[..., :48], g[..., 48:]
    g = layer_back("d3", up_backward(g_up3))
    g_up4, g_s3 = g[..., :64], g[..., 64:]
    g = layer_back("e4", up_backward(g_up4))
    g_s3 = g_s3 + pool_backward(g)
    g = layer_back("e3", g_s3)
    g_s2 = g_s2 + pool_backward(g)
    g = layer_back("e2", g_s2)
    g_s1 = g_s1 + pool_backward(g)
    layer_back("e1", g_s1)
    return G


# ---------------------------------------------------------------------------
# energy-norm loss: the share of the error's A-norm that the exact line
# search along the upsampled output removes
# ---------------------------------------------------------------------------
def up_matrix(nf, nc, pool=4):
    """1-D cell-centred linear interpolation from nc coarse to nf fine
    cells: fine x sits at coarse coordinate (x + 0.5)/pool - 0.5, clamped to
    [0, nc - 1]."""
    M = np.zeros((nf, nc))
    for x in range(nf):
        c = min(max((x + 0.5) / pool - 0.5, 0.0), nc - 1.0)
        x0 = min(int(np.floor(c)), nc - 1)
        x1 = min(x0 + 1, nc - 1)
        f = c - x0
        M[x, x0] += 1.0 - f
        M[x, x1] += f
    return M


def upsample(y, fixed):
    """y (nx/4, ny/4, nz/4) -> (nx, ny, nz), 0 on the fixed cells."""
    nx, ny, nz = fixed.shape
    Mx, My, Mz = (up_matrix(n, c) for n, c in zip((nx, ny, nz), y.shape))
    u = np.einsum("ia,jb,kc,abc->ijk", Mx, My, Mz, y)
    u[fixed] = 0.0
    return u


def upsample_adjoint(g, fixed, coarse_shape):
    nx, ny, nz = fixed.shape
    Mx, My, Mz = (up_matrix(n, c) for n, c in zip((nx, ny, nz), coarse_shape))
    return np.einsum("ia,jb,kc,ijk->abc", Mx, My, Mz, np.where(fixed, 0.0, g))


def apply_op(faces, fixed, v):
    """7-point operator: faces (3, nx, ny, nz) hold each cell's +x/+y/+z face
    (0 on the global edge); free rows d v - sum t v_nb with d the sum of the
    cell's six faces, fixed rows 0."""
    tx, ty, tz = faces
    d = tx + ty + tz
    d[1:, :, :] += tx[:-1, :, :]
    d[:, 1:, :] += ty[:, :-1, :]
    d[:, :, 1:] += tz[:, :, :-1]
    s = np.zeros_like(v)
    s[:-1, :, :] += tx[:-1, :, :] * v[1:, :, :]
    s[1:, :, :] += tx[:-1, :, :] * v[:-1, :, :]
    s[:, :-1, :] += ty[:, :-1, :] * v[:, 1:, :]
    s[:, 1:, :] += ty[:, :-1, :] * v[:, :-1, :]
    s[:, :, :-1] += tz[:, :, :-1] * v[:, :, 1:]
    s[:, :, 1:] += tz[:, :, :-1] * v[:, :, :-1]
    out = d * v - s
    out[fixed] = 0.0
    return out


def energy_loss(y, target, faces, fixed):
    """L = 1 - (u, At)^2 / ((u, Au)(t, At)) for u = upsample(y); returns L
    and dL/dy with y's shape (..., 1)."""
    yc = y[..., 0]
    u = upsample(yc, fixed)
    At, Au = apply_op(faces, fixed, target), apply_op(faces, fixed, u)
    a, b, c = float((u * At).sum()), float((u * Au).sum()), float((target * At).sum())
    L = 1.0 - a * a / (b * c)
    gu = -2.0 * a / (b * c) * At + 2.0 * a * a / (b * b * c) * Au
    return L, upsample_adjoint(gu, fixed, yc.shape)[..., None]
