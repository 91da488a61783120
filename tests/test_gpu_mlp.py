"""GPU checks of the learned in-scattering network's hand-written MFMA kernels
against the FP32 NumPy oracle (oracle/oracle_mlp.py).

The kernels run bf16 weights and activations with fp32 accumulation; the
oracle is given the same bf16-rounded weights and inputs, so the remaining
difference is the bf16 rounding of hidden activations (a few 1e-3 relative).
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def M():
    spec = importlib.util.spec_from_file_location("oracle_mlp", os.path.join(ROOT, "oracle", "oracle_mlp.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def bf16(a):
    """float32 -> bf16 bits (round to nearest even) and the rounded float32 values."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r, (r.astype(np.uint32) << 16).view(np.float32)


def layout(pkg):
    K, N, w, b, t = (C.c_uint32 * 5)(), (C.c_uint32 * 5)(), (C.c_uint32 * 5)(), (C.c_uint32 * 5)(), C.c_uint32()
    pkg.load().cr_mlp_param_layout(K, N, w, b, C.byref(t))
    return list(K), list(N), list(w), list(b), t.value


def pack(pkg, params):
    """Oracle params (real shapes) -> padded flat fp32 layout."""
    K, N, wo, bo, total = layout(pkg)
    flat = np.zeros(total, np.float32)
    for l in range(5):
        w, b = params[2 * l], params[2 * l + 1]
        W = np.zeros((K[l], N[l]), np.float32)
        W[: w.shape[0], : w.shape[1]] = w
        flat[wo[l]: wo[l] + K[l] * N[l]] = W.ravel()
        flat[bo[l]: bo[l] + b.shape[0]] = b
    return flat


def unpack(pkg, flat, shapes):
    K, N, wo, bo, total = layout(pkg)
    out = []
    for l, (i, o) in enumerate(shapes):
        W = flat[wo[l]: wo[l] + K[l] * N[l]].reshape(K[l], N[l])[:i, :o]
        out += [W.copy(), flat[bo[l]: bo[l] + o].copy()]
    return out


def test_mfma_lane_layout(ctx, pkg):
    """One v_mfma_f32_16x16x32_bf16 with the operand layout the kernels assume
    (A[m][8g + j], B[8g + j][m], D[4g + j][m] for lane 16g + m) equals the
    host product of the same bf16 matrices."""
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 32)).astype(np.float32)
    B = rng.standard_normal((32, 16)).astype(np.float32)
    Ab, Af = bf16(A)
    Bb, Bf = bf16(B)
    D = np.zeros((16, 16), np.float32)
    ctx.check(pkg.load().cr_mlp_probe(ctx.h, Ab.ctypes.data, Bb.ctypes.data, D.ctypes.data), "probe")
    np.testing.assert_allclose(D, Af.astype(np.float64) @ Bf.astype(np.float64), rtol=1e-5, atol=1e-5)


def make_params(M, seed):
    ps = M.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for l in range(5):
        ps[2 * l + 1] = (rng.standard_normal(ps[2 * l + 1].shape) * 0.05).astype(np.float32)
    return [bf16(p)[1] if i % 2 == 0 else p for i, p in enumerate(ps)]  # weights exactly bf16


def test_forward_matches_oracle(ctx, pkg, M):
    ps = make_params(M, 1)
    flat = pack(pkg, ps)
    ctx.check(pkg.load().cr_mlp_set_params(ctx.h, flat.ctypes.data, len(flat)), "set params")
    B = 300  # not a multiple of the 128-row tile
    X = np.random.default_rng(2).random((B, 224)).astype(np.float32)
    Xb, Xf = bf16(X)
    Y = np.zeros((B, 3), np.float32)
    ctx.check(pkg.load().cr_mlp_forward(ctx.h, Xb.ctypes.data, B, Y.ctypes.data), "forward")
    ref, _ = M.forward(ps, Xf)
    err = np.abs(Y - ref).max() / (np.abs(ref).max() + 1e-6)
    assert err < 2e-2, err


# 256 rows: 8 tiles of 32 rows and one gradient slot; 70000 rows: 2188 tiles, the last one ragged,
# and 16 gradient slots (the most the weight-gradient GEMM uses)
@pytest.mark.parametrize("B", [256, 70000])
def test_gradients_match_oracle(ctx, pkg, M, B):
    ps = make_params(M, 3)
    flat = pack(pkg, ps)
    lib = pkg.load()
    ctx.check(lib.cr_mlp_set_params(ctx.h, flat.ctypes.data, len(flat)), "set params")
    rng = np.random.default_rng(4)
    X = rng.random((B, 224)).astype(np.float32)
    T = rng.random((B, 3)).astype(np.float32) * 3.0
    Xb, Xf = bf16(X)
    loss = C.c_float()
    grads = np.zeros(len(flat), np.float32)
    ctx.check(lib.cr_mlp_train_step(ctx.h, Xb.ctypes.data, T.ctypes.data, B, 0.0, C.byref(loss), grads.ctypes.data),
              "train step")
    ref_loss, ref_g = M.loss_and_grads(ps, Xf, T)
    assert abs(loss.value - ref_loss) < 2e-2 * ref_loss + 1e-6
    got = unpack(pkg, grads, M.LAYERS)
    for i, (g, r) in enumerate(zip(got, ref_g)):
        rel = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-12)
        assert rel < 5e-2, (i, rel)
    # padding of the 3-wide output layer stays untouched
    K, N, wo, bo, total = layout(pkg)
    pad = grads[wo[4]: wo[4] + K[4] * N[4]].reshape(K[4], N[4])[:, 3:]
    assert not pad.any()


def test_training_is_bit_reproducible(ctx, pkg, M):
    """The weight-gradient GEMM sums each gradient slot in a fixed order and
    the Adam kernel adds the slots in order (no float atomics), so the same
    steps from the same parameters give bit-identical gradients, losses and
    parameters."""
    ps = make_params(M, 9)
    flat = pack(pkg, ps)
    lib = pkg.load()
    B = 5000
    rng = np.random.default_rng(10)
    Xb, _ = bf16(rng.random((B, 224)).astype(np.float32))
    T = (rng.random((B, 3)) * 2.0).astype(np.float32)
    runs = []
    for _ in range(2):
        ctx.check(lib.cr_mlp_set_params(ctx.h, flat.ctypes.data, len(flat)), "set params")
        losses, grads = [], np.zeros(len(flat), np.float32)
        loss = C.c_float()
        for _ in range(5):
            ctx.check(lib.cr_mlp_train_step(ctx.h, Xb.ctypes.data, T.ctypes.data, B, 1e-3, C.byref(loss),
                                            grads.ctypes.data), "train step")
            losses.append(loss.value)
        after = np.zeros(len(flat), np.float32)
        ctx.check(lib.cr_mlp_get_params(ctx.h, after.ctypes.data, len(after)), "get params")
        runs.append((np.array(losses, np.float32), grads.copy(), after))
    for (a, b, what) in zip(runs[0], runs[1], ("losses", "gradients", "parameters")):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), what
    assert not np.array_equal(runs[0][2], flat)


def test_parameter_moving_average(ctx, pkg, M):
    """cr_mlp_set_ema: after each Adam step the device holds d * ema + (1 - d)
    * params, bias-corrected by 1 / (1 - d^t) (float32 arithmetic, checked bit
    for bit against the same sequence on the host), and inference uses it: the
    forward pass equals the forward pass of a network whose parameters are set
    to the average.  Decay 0 turns it off again."""
    ps = make_params(M, 11)
    flat = pack(pkg, ps)
    lib = pkg.load()
    total = len(flat)
    ctx.check(lib.cr_mlp_set_params(ctx.h, flat.ctypes.data, total), "set params")
    d = 0.8
    ctx.check(lib.cr_mlp_set_ema(ctx.h, d), "set ema")
    B = 512
    rng = np.random.default_rng(12)
    Xb, _ = bf16(rng.random((B, 224)).astype(np.float32))
    T = (rng.random((B, 3)) * 2.0).astype(np.float32)
    loss = C.c_float()
    e = np.zeros(total, np.float32)
    p = np.zeros(total, np.float32)
    for t in range(1, 4):
        ctx.check(lib.cr_mlp_train_step(ctx.h, Xb.ctypes.data, T.ctypes.data, B, 0.02, C.byref(loss), None), "step")
        ctx.check(lib.cr_mlp_get_params(ctx.h, p.ctypes.data, total), "get params")
        e = np.float32(d) * e + (np.float32(1.0) - np.float32(d)) * p
        corr = np.float32(1.0 / (1.0 - float(np.float32(d)) ** t))
        got = np.zeros(total, np.float32)
        ctx.check(lib.cr_mlp_get_ema(ctx.h, got.ctypes.data, total), "get ema")
        assert np.array_equal((e * corr).view(np.uint32), got.view(np.uint32)), f"step {t}"
    assert not np.array_equal(got, p)
    Y_ema = np.zeros((B, 3), np.float32)
    ctx.check(lib.cr_mlp_forward(ctx.h, Xb.ctypes.data, B, Y_ema.ctypes.data), "forward (average)")
    ctx.check(lib.cr_mlp_set_ema(ctx.h, 0.0), "ema off")
    Y_raw = np.zeros((B, 3), np.float32)
    ctx.check(lib.cr_mlp_forward(ctx.h, Xb.ctypes.data, B, Y_raw.ctypes.data), "forward (parameters)")
    ctx.check(lib.cr_mlp_set_params(ctx.h, got.ctypes.data, total), "set params to the average")
    Y_set = np.zeros((B, 3), np.float32)
    ctx.check(lib.cr_mlp_forward(ctx.h, Xb.ctypes.data, B, Y_set.ctypes.data), "forward (set)")
    assert np.array_equal(Y_ema.view(np.uint32), Y_set.view(np.uint32))
    assert not np.array_equal(Y_ema, Y_raw)
    assert lib.cr_mlp_set_ema(ctx.h, 1.0) != 0  # decay must be < 1


def test_adam_step_matches_oracle(ctx, pkg, M):
    ps = make_params(M, 5)
    flat = pack(pkg, ps)
    lib = pkg.load()
    ctx.check(lib.cr_mlp_set_params(ctx.h, flat.ctypes.data, len(flat)), "set params")
    B = 128
    rng = np.random.default_rng(6)
    Xb, Xf = bf16(rng.random((B, 224)).astype(np.float32))
    T = rng.random((B, 3)).astype(np.float32)
    grads = np.zeros(len(flat), np.float32)
    loss = C.c_float()
    ctx.check(lib.cr_mlp_train_step(ctx.h, Xb.ctypes.data, T.ctypes.data, B, 1e-3, C.byref(loss), grads.ctypes.data),
              "train step")
    after = np.zeros(len(flat), np.float32)
    ctx.check(lib.cr_mlp_get_params(ctx.h, after.ctypes.data, len(after)), "get params")
    # Adam step 1 from zero moments, applied by the oracle to the GPU's own gradients
    ref = [p.copy() for p in ps]
    g_own = unpack(pkg, grads, M.LAYERS)
    M.adam_step(ref, g_own, [np.zeros_like(p) for p in ps], [np.zeros_like(p) for p in ps], 1, lr=1e-3)
    got = unpack(pkg, after, M.LAYERS)
    for r, g in zip(ref, got):
        np.testing.assert_allclose(g, r, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(ctx, pkg, M):
    """A smooth target the network can fit: the loss drops by 10x in 300 steps."""
    lib = pkg.load()
    ctx.check(lib.cr_mlp_init(ctx.h, 7), "init")
    B = 1024
    rng = np.random.default_rng(8)
    X = rng.random((B, 224)).astype(np.float32)
    T = np.expm1(np.stack([X[:, :8].mean(1), X[:, 8:16].mean(1) * 2, X[:, 216] ** 2], 1)).astype(np.float32)
    Xb, _ = bf16(X)
    loss = C.c_float()
    losses = []
    for it in range(300):
        ctx.check(lib.cr_mlp_train_step(ctx.h, Xb.ctypes.data, T.ctypes.data, B, 1e-3, C.byref(loss), None), "step")
        losses.append(loss.value)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
