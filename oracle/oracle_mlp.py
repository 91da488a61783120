"""ORACLE — test infrastructure only.

FP32 NumPy restatement of the learned in-scattering network (an extension:
the reference has no learned component).  Used by tests/ to check the gfx950
kernels' features, forward pass, gradients and optimiser step.

Network: 224 -> 128 -> 128 -> 128 -> 128 -> 3, ReLU after each hidden layer,
linear output = log(1 + Y) per RGB channel, where Y is the radiance the
plain path tracer adds after a real collision at x for a camera path arriving
along omega (the in-scattered radiance, RPNN: Kallweit et al., "Deep
Scattering", 2017).  Weights are stored in row-major [in, out] order.

Features (224): for mip levels k = 0..7 of the density grid, 27 samples on a
3x3x3 stencil around x with spacing h_k = h_0 2^k in the frame whose z axis
points to the sun (-L) and whose x, y axes are the Duff et al. basis of z;
each sample is the optical depth density * (densityScaling/baseScaling) * h_k
(0 outside the box), then cos(omega, -L), its square and six zeros.
"""
import numpy as np

LAYERS = [(224, 128), (128, 128), (128, 128), (128, 128), (128, 3)]
N_PARAMS = sum(i * o + o for i, o in LAYERS)  # 78,723
FLOPS_FWD = sum(2 * i * o for i, o in LAYERS)  # 156,416
LEVELS = 8


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    ps = []
    for i, o in LAYERS:
        w = (rng.standard_normal((i, o)) * np.sqrt(2.0 / i)).astype(np.float32)
        b = np.zeros(o, np.float32)
        ps += [w, b]
    return ps


def forward(params, x):
    """Returns the output and the cached activations (post-ReLU inputs of each layer)."""
    acts = [x.astype(np.float32)]
    h = acts[0]
    for li in range(len(LAYERS)):
        w, b = params[2 * li], params[2 * li + 1]
        z = h @ w + b
        if li < len(LAYERS) - 1:
            h = np.maximum(z, 0.0).astype(np.float32)
            acts.append(h)
        else:
            h = z.astype(np.float32)
    return h, acts


def loss_and_grads(params, x, target):
    """Mean squared error of the output against log(1 + target); gradients of
    every parameter (FP32)."""
    y, acts = forward(params, x)
    t = np.log1p(np.maximum(target, 0.0)).astype(np.float32)
    n = x.shape[0]
    diff = y - t
    loss = float((diff * diff).mean())
    g = (2.0 / (n * y.shape[1])) * diff
    grads = [None] * len(params)
    for li in reversed(range(len(LAYERS))):
        w = params[2 * li]
        a = acts[li]
        grads[2 * li] = (a.T @ g).astype(np.float32)
        grads[2 * li + 1] = g.sum(axis=0).astype(np.float32)
        if li > 0:
            g = (g @ w.T) * (acts[li] > 0)
    return loss, grads


def adam_step(params, grads, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """In-place Adam (Kingma & Ba 2015) with bias correction."""
    c1 = 1.0 - b1 ** step
    c2 = 1.0 - b2 ** step
    for p, g, mm, vv in zip(params, grads, m, v):
        mm *= b1
        mm += (1.0 - b1) * g
        vv *= b2
        vv += (1.0 - b2) * g * g
        p -= lr * (mm / c1) / (np.sqrt(vv / c2) + eps)


def mip_pyramid(grid, levels=LEVELS):
    """Level k+1 is the mean of 2x2x2 blocks of level k (edge blocks average
    the voxels that exist)."""
    mips = [grid.astype(np.float32)]
    for _ in range(1, levels):
        g = mips[-1]
        nz, ny, nx = g.shape
        oz, oy, ox = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
        acc = np.zeros((oz, oy, ox), np.float32)
        cnt = np.zeros((oz, oy, ox), np.float32)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    s = g[dz::2, dy::2, dx::2]
                    acc[: s.shape[0], : s.shape[1], : s.shape[2]] += s
                    cnt[: s.shape[0], : s.shape[1], : s.shape[2]] += 1.0
        mips.append(acc / cnt)
    return mips


def onb(n):
    sz = 1.0 if n[2] >= 0 else -1.0
    a = n[1] / (1.0 + abs(n[2]))
    b = n[1] * a
    c = -n[0] * a
    return np.array([n[2] + sz * b, sz * c, -n[0]]), np.array([c, 1.0 - b, -sz * n[1]])


def trilinear(g, n):
    """Linear clamp-to-edge lookup at normalised coordinates n (x, y, z)."""
    nz, ny, nx = g.shape
    out = 0.0
    idx, w = [], []
    for c, N in zip(n, (nx, ny, nz)):
        u = min(max(c * N - 0.5, -1.0), N)
        f = np.floor(u)
        a = u - f
        i0 = int(min(max(f, 0), N - 1))
        i1 = int(min(max(f + 1, 0), N - 1))
        idx.append((i0, i1))
        w.append(a)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wt = (w[0] if dx else 1 - w[0]) * (w[1] if dy else 1 - w[1]) * (w[2] if dz else 1 - w[2])
                out += wt * g[idx[2][dz], idx[1][dy], idx[0][dx]]
    return out


def features(mips, b0, b1, sigma_scale, h0, x, omega, light_dir):
    """224 features of one shading point (see the module docstring)."""
    b0 = np.asarray(b0, np.float64)[:3]
    b1 = np.asarray(b1, np.float64)[:3]
    ez = -np.asarray(light_dir, np.float64)[:3]
    ex, ey = onb(ez)
    f = np.zeros(224, np.float32)
    j = 0
    for k in range(LEVELS):
        h = h0 * (2.0 ** k)
        for l in (-1, 0, 1):
            for jj in (-1, 0, 1):
                for i in (-1, 0, 1):
                    p = np.asarray(x, np.float64) + h * (i * ex + jj * ey + l * ez)
                    if np.all(p >= b0) and np.all(p <= b1):
                        d = trilinear(mips[k], (p - b0) / (b1 - b0))
                    else:
                        d = 0.0
                    f[j] = d * sigma_scale * h
                    j += 1
    c = float(np.dot(np.asarray(omega, np.float64), ez))
    f[216] = c
    f[217] = c * c
    return f
