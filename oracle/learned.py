"""Oracle of the learned convolutional preconditioner (test infrastructure only).

The reference has no learned preconditioner: the network is this project's addition to the
ConjugateGradientPoissonSolver path (conjugate_gradient_poisson_solver.jl:169-211, where the FFT
preconditioner is applied).  This module restates, in NumPy, exactly what the HIP kernels compute,
so the GPU forward pass, its gradients and a training step can be checked:

* four 3×3×3 convolutions with 16 input and 16 output channels each (cross-correlation,
  y[v, o] = Σ_tap Σ_c x[v + tap, c] · W[tap, c, o]), ReLU after the first three;
* taps ordered t = (dz + 1)·9 + (dy + 1)·3 + (dx + 1); x and y periodic, z zero-padded;
* the network input has 3 meaningful channels (scaled residual, scaled FFT-preconditioned
  residual, active-cell mask) padded with zeros to 16, and the output is channel 0 of layer 4;
* activations, weights and back-propagated gradients are rounded to BF16 (round to nearest even)
  where the device stores them (they are MFMA operands); products and sums are FP32
  (`precise=False`) or FP64 without rounding (`precise=True`, for finite differences).

Loss: L = Σ_active (y − t)² / N_active over the active cells (mask = 1).
"""
import numpy as np

NL, K3, C = 4, 27, 16


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


def pad_input(x3):
    """(Nx, Ny, Nz, 3) → (Nx, Ny, Nz, 16) with zero channels 3..15."""
    x = np.zeros(x3.shape[:3] + (C,), dtype=np.result_type(x3, np.float32))
    x[..., :x3.shape[3]] = x3
    return x


def forward(x3, weights, precise=False):
    """Network output (Nx, Ny, Nz) and the stored activations [x0, y1, y2, y3] (BF16 values)."""
    dt = np.float64 if precise else np.float32
    r = (lambda a: np.asarray(a, dtype=np.float64)) if precise else bf16
    w = [r(weights[l]) for l in range(NL)]
    acts = [r(pad_input(x3))]
    h = acts[0]
    for l in range(NL):
        y = conv(h, w[l], dt)
        if l < NL - 1:
            h = r(np.maximum(y, 0))
            acts.append(h)
        else:
            out = y[..., 0].astype(dt)
    return out, acts


def loss_and_grad_output(out, target, mask):
    n = max(float(mask.sum()), 1.0)
    d = (out - target) * mask
    return float((d * d).sum() / n), (2.0 / n) * d


def backward(acts, weights, dout, precise=False):
    """Gradients of L w.r.t. the weights (4, 27, 16, 16) and the padded input (Nx, Ny, Nz, 16),
    given dL/d(out).  Like the device, the gradient w.r.t. each layer's output is stored in BF16
    (it is an MFMA operand) and every product is accumulated in FP32."""
    dt = np.float64 if precise else np.float32
    r = (lambda a: np.asarray(a, dtype=np.float64)) if precise else bf16
    w = [r(weights[l]) for l in range(NL)]
    g = np.zeros(dout.shape + (C,), dtype=np.float32 if not precise else np.float64)
    g[..., 0] = dout
    g = r(g)
    dws = [None] * NL
    for l in range(NL - 1, -1, -1):
        dws[l] = weight_grad(acts[l], g, dt)
        dx = conv_transpose(g, w[l], dt)
        if l > 0:
            g = r(dx * (acts[l] > 0))
        else:
            gin = dx
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
