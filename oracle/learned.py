"""Oracle of the learned convolutional preconditioner (test infrastructure only).

The reference has no learned preconditioner: the network is this project's addition to the
ConjugateGradientPoissonSolver path (conjugate_gradient_poisson_solver.jl:169-211, where the FFT
preconditioner is applied).  This module restates, in NumPy, exactly what the HIP kernels compute,
so the GPU forward pass, its gradients and a training step can be checked.

The network is a four-level hierarchy of 3×3×3 convolutions with 16 input and 16 output channels
each (cross-correlation, y[v, o] = Σ_tap Σ_c x[v + tap, c] · W[tap, c, o]; taps ordered
t = (dz + 1)·9 + (dy + 1)·3 + (dx + 1); x and y periodic, z zero-padded), seven layers:

    a0 = relu(conv0(x))              level 0 (the grid)
    a1 = relu(conv1(pool(a0)))       level 1 (2× coarser in x, y and z)
    a2 = relu(conv2(pool(a1)))       level 2
    a3 = relu(conv3(pool(a2)))       level 3
    b2 = relu(conv4(a2 + up(a3)))    level 2
    b1 = relu(conv5(a1 + up(b2)))    level 1
    out = conv6(a0 + up(b1))[0]      level 0, output channel 0

pool: the mean of the 2×2×2 children, summed in the order (z, y, x) of the child offsets; up:
each coarse value copied to its 8 children.  The input has 3 meaningful channels (scaled residual,
scaled FFT-preconditioned residual, active-cell mask) padded with zeros to 16.  Activations,
pooled and summed values, weights and back-propagated gradients are rounded to BF16 (round to
nearest even) where the device stores them; products and sums are FP32 (`precise=False`) or FP64
without rounding (`precise=True`, for finite differences).

Loss: L = Σ_active (y − t)² / N_active over the active cells (mask = 1).
"""
import numpy as np

NL, K3, C = 7, 27, 16
LEVEL = (0, 1, 2, 3, 2, 1, 0)   # the level each layer runs on


def bf16(a):
    """Round FP32 values to BF16 (nearest, ties to even) and return them as FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def taps():
    return [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(x, dx, dy, dz):
    """s[i, j, k] = x[i + dx, j + dy, k + dz]: periodic in x and y, zero outside in z.
    x has shape (Nx, Ny, Nz, ...)."""
    s = np.roll(x, (-dx, -dy), axis=(0, 1))
    if dz == 0:
        return s
    out = np.zeros_like(s)
    if dz > 0:
        out[:, :, :-dz] = s[:, :, dz:]
    else:
        out[:, :, -dz:] = s[:, :, :dz]
    return out


def conv(x, w, dtype=np.float32):
    """y[v, o] = Σ_t Σ_c x[v + tap_t, c] w[t, c, o]; x (Nx, Ny, Nz, 16), w (27, 16, 16)."""
    y = np.zeros(x.shape[:3] + (w.shape[2],), dtype=dtype)
    for t, (dz, dy, dx) in enumerate(taps()):
        y += shift(x, dx, dy, dz).astype(dtype) @ w[t].astype(dtype)
    return y


def conv_transpose(g, w, dtype=np.float32):
    """Adjoint of conv in x: dx[v, c] = Σ_t Σ_o g[v − tap_t, o] w[t, c, o]."""
    dx = np.zeros(g.shape[:3] + (w.shape[1],), dtype=dtype)
    for t, (dz, dy, dxx) in enumerate(taps()):
        dx += shift(g, -dxx, -dy, -dz).astype(dtype) @ w[t].T.astype(dtype)
    return dx


def weight_grad(x, g, dtype=np.float32):
    """dW[t, c, o] = Σ_v x[v + tap_t, c] g[v, o]."""
    dw = np.zeros((K3, x.shape[3], g.shape[3]), dtype=dtype)
    G = g.reshape(-1, g.shape[3]).astype(dtype)
    for t, (dz, dy, dx) in enumerate(taps()):
        X = shift(x, dx, dy, dz).reshape(-1, x.shape[3]).astype(dtype)
        dw[t] = X.T @ G
    return dw


def child_sum(x, dtype=np.float32):
    """Σ of the 2×2×2 children of every coarse cell, added in the order (z, y, x) of the offsets."""
    s = np.zeros((x.shape[0] // 2, x.shape[1] // 2, x.shape[2] // 2) + x.shape[3:], dtype=dtype)
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                s = (s + x[cx::2, cy::2, cz::2].astype(dtype)).astype(dtype)
    return s


def pool(x, dtype=np.float32):
    return (child_sum(x, dtype) * dtype(0.125)).astype(dtype)


def up(x):
    """Each coarse value copied to its 8 children."""
    return np.repeat(np.repeat(np.repeat(x, 2, axis=0), 2, axis=1), 2, axis=2)


def pad_input(x3):
    """(Nx, Ny, Nz, 3) → (Nx, Ny, Nz, 16) with zero channels 3..15."""
    x = np.zeros(x3.shape[:3] + (C,), dtype=np.result_type(x3, np.float32))
    x[..., :x3.shape[3]] = x3
    return x


def check_dims(shape):
    nx, ny, nz = shape[:3]
    if nx % 128 or ny % 8 or nz % 8:
        raise ValueError("the four-level network needs Nx a multiple of 128 and Ny, Nz multiples of 8")


def forward(x3, weights, precise=False):
    """Network output (Nx, Ny, Nz) and the stored tensors the backward pass reads."""
    check_dims(x3.shape)
    dt = np.float64 if precise else np.float32
    r = (lambda a: np.asarray(a, dtype=np.float64)) if precise else bf16
    w = [r(weights[l]) for l in range(NL)]
    relu = lambda a: np.maximum(a, 0)
    st = {}
    st["x0"] = r(pad_input(x3))
    st["a0"] = r(relu(conv(st["x0"], w[0], dt)))
    st["p1"] = r(pool(st["a0"], dt))
    st["a1"] = r(relu(conv(st["p1"], w[1], dt)))
    st["p2"] = r(pool(st["a1"], dt))
    st["a2"] = r(relu(conv(st["p2"], w[2], dt)))
    st["p3"] = r(pool(st["a2"], dt))
    st["a3"] = r(relu(conv(st["p3"], w[3], dt)))
    st["s2"] = r(st["a2"].astype(dt) + up(st["a3"]).astype(dt))
    st["b2"] = r(relu(conv(st["s2"], w[4], dt)))
    st["s1"] = r(st["a1"].astype(dt) + up(st["b2"]).astype(dt))
    st["b1"] = r(relu(conv(st["s1"], w[5], dt)))
    st["s0"] = r(st["a0"].astype(dt) + up(st["b1"]).astype(dt))
    out = conv(st["s0"], w[6], dt)[..., 0].astype(dt)
    return out, st


def loss_and_grad_output(out, target, mask):
    n = max(float(mask.sum()), 1.0)
    d = (out - target) * mask
    return float((d * d).sum() / n), (2.0 / n) * d


def backward(st, weights, dout, precise=False):
    """Gradients of L w.r.t. the weights (7, 27, 16, 16) and the padded input (Nx, Ny, Nz, 16),
    given dL/d(out).  Like the device, every back-propagated gradient that is stored (an MFMA
    operand) is rounded to BF16 and every product is accumulated in FP32."""
    dt = np.float64 if precise else np.float32
    r = (lambda a: np.asarray(a, dtype=np.float64)) if precise else bf16
    w = [r(weights[l]) for l in range(NL)]
    pos = lambda a: (a > 0)
    g6 = np.zeros(dout.shape + (C,), dtype=dt)
    g6[..., 0] = dout
    g6 = r(g6)
    dws = [None] * NL
    dws[6] = weight_grad(st["s0"], g6, dt)
    gs0 = r(conv_transpose(g6, w[6], dt))
    g5 = r(child_sum(gs0, dt) * pos(st["b1"]))
    dws[5] = weight_grad(st["s1"], g5, dt)
    gs1 = r(conv_transpose(g5, w[5], dt))
    g4 = r(child_sum(gs1, dt) * pos(st["b2"]))
    dws[4] = weight_grad(st["s2"], g4, dt)
    gs2 = r(conv_transpose(g4, w[4], dt))
    g3 = r(child_sum(gs2, dt) * pos(st["a3"]))
    dws[3] = weight_grad(st["p3"], g3, dt)
    gp3 = r(conv_transpose(g3, w[3], dt))
    g2 = r((gs2.astype(dt) + dt(0.125) * up(gp3).astype(dt)) * pos(st["a2"]))
    dws[2] = weight_grad(st["p2"], g2, dt)
    gp2 = r(conv_transpose(g2, w[2], dt))
    g1 = r((gs1.astype(dt) + dt(0.125) * up(gp2).astype(dt)) * pos(st["a1"]))
    dws[1] = weight_grad(st["p1"], g1, dt)
    gp1 = r(conv_transpose(g1, w[1], dt))
    g0 = r((gs0.astype(dt) + dt(0.125) * up(gp1).astype(dt)) * pos(st["a0"]))
    dws[0] = weight_grad(st["x0"], g0, dt)
    gin = conv_transpose(g0, w[0], dt)
    return np.stack(dws), gin


def energy(d, dx, dy, dzc, dzf, active):
    """The training loss's energy E = dᵀ(−V∇²)d of a masked error d (Nx, Ny, Nz), with the CG
    solver's conditional operator: a sum over the faces between two active cells of
    (face area / spacing)·(d_a − d_b)², x and y periodic, no flux through the z walls or into
    inactive cells.  dzc[k] are the cell heights, dzf[k] (k = 1 .. Nz−1) the spacings between
    cell centres k−1 and k.  Returns (E, dE/dd) in FP64."""
    d = np.where(active, np.asarray(d, dtype=np.float64), 0.0)
    Nx, Ny, Nz = d.shape
    E = 0.0
    g = np.zeros_like(d)
    dzc = np.asarray(dzc, dtype=np.float64)
    for axis, coef in ((0, dy * dzc / dx), (1, dx * dzc / dy)):
        nb = np.roll(d, 1, axis=axis)
        both = active & np.roll(active, 1, axis=axis)
        c = np.where(both, np.broadcast_to(coef[None, None, :], d.shape), 0.0)
        diff = d - nb
        E += float((c * diff * diff).sum())
        g += 2 * c * diff
        g -= np.roll(2 * c * diff, -1, axis=axis)
    for k in range(1, Nz):
        c = np.where(active[:, :, k] & active[:, :, k - 1], dx * dy / dzf[k], 0.0)
        diff = d[:, :, k] - d[:, :, k - 1]
        E += float((c * diff * diff).sum())
        g[:, :, k] += 2 * c * diff
        g[:, :, k - 1] -= 2 * c * diff
    return E, np.where(active, g, 0.0)


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in FP32, as the device applies it."""
    p, g = params.astype(np.float32), grads.astype(np.float32)
    m = (np.float32(beta1) * m + np.float32(1 - beta1) * g).astype(np.float32)
    v = (np.float32(beta2) * v + np.float32(1 - beta2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - beta1 ** t)
    vh = v / np.float32(1 - beta2 ** t)
    return (p - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32), m, v


def init_weights(seed=0, scale=None):
    """He-style initialisation; the output layer starts small so the preconditioner starts near
    the FFT preconditioner."""
    rng = np.random.default_rng(seed)
    w = np.zeros((NL, K3, C, C), dtype=np.float32)
    fan_in = K3 * C
    for l in range(NL):
        w[l] = rng.normal(0, np.sqrt(2.0 / fan_in), size=(K3, C, C)).astype(np.float32)
    w[0, :, 3:, :] = 0        # padded input channels
    w[NL - 1, :, :, 1:] = 0   # only output channel 0 is used
    w[NL - 1] *= 0.1 if scale is None else scale
    return w
