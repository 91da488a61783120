"""The kernels' elementary functions (sqrt, log, exp, sin, cos) on the device
against the oracle's CPU implementation, bit for bit, over every binade of the
inputs the path tracer and the grid generator feed them.

sqrt is also checked against NumPy's float32 square root, which IEEE-754
requires to be correctly rounded: the hardware square root alone is accurate to
one ULP, so this is what pins the device's rounding correction.
"""
import numpy as np
import pytest

from gpu_util import assert_bitexact

pytestmark = pytest.mark.gpu

SQRT, LOG, EXP, SIN, COS = 0, 1, 2, 3, 4


def probe(pkg, ctx, op, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    ctx.check(pkg.load().cr_math_probe(ctx.h, op, x.ctypes.data, x.size, out.ctypes.data), "math probe")
    return out


def positive_floats(rng, n):
    """Uniform over the bit patterns of positive finite floats (denormals included)."""
    bits = rng.integers(1, 0x7F800000, size=n, dtype=np.uint32)
    return bits.view(np.float32)


def specials():
    return np.array([0.0, -0.0, 1.0, 2.0, 4.0, 0.25, np.inf, -np.inf, np.nan, -1.0, 1e-45, 1.17549435e-38,
                     3.4028235e38, 0.5, 0.999999940, 1.00000012], np.float32)


def test_sqrt_correctly_rounded(pkg, ctx):
    rng = np.random.default_rng(1)
    x = np.concatenate([positive_floats(rng, 1 << 20), rng.random(1 << 18, dtype=np.float32), specials()])
    got = probe(pkg, ctx, SQRT, x)
    with np.errstate(invalid="ignore"):
        ref = np.sqrt(x)
    assert_bitexact(got, ref, "device sqrt vs IEEE sqrt")


@pytest.mark.parametrize("op", [LOG, EXP, SIN, COS])
def test_elementary_functions_bitexact(pkg, ctx, O, op):
    rng = np.random.default_rng(10 + op)
    if op == LOG:
        x = np.concatenate([positive_floats(rng, 1 << 20), rng.random(1 << 18, dtype=np.float32)])
    elif op == EXP:
        x = rng.uniform(-104.0, 89.0, 1 << 20).astype(np.float32)
    else:
        x = np.concatenate([rng.uniform(-7.0, 7.0, 1 << 20), rng.uniform(-200.0, 200.0, 1 << 16)]).astype(np.float32)
    x = np.concatenate([x, specials()])
    assert_bitexact(probe(pkg, ctx, op, x), O.math_array(op, x), f"device vs oracle, op {op}")


def test_probe_rejects_unknown_op(pkg, ctx):
    x = np.ones(4, np.float32)
    out = np.empty_like(x)
    assert pkg.load().cr_math_probe(ctx.h, 9, x.ctypes.data, 4, out.ctypes.data) != 0


@pytest.mark.parametrize("op", [5, 6])
def test_inverse_trig_bitexact(pkg, ctx, O, op):
    """acos and atan2 (the photon map's direction angles) on the device
    against the oracle, bit for bit; atan2 pairs neighbouring inputs."""
    rng = np.random.default_rng(30 + op)
    if op == 5:
        x = np.concatenate([rng.uniform(-1.0, 1.0, 1 << 20), [-1.0, 1.0, 0.0, -0.0, 0.5, -0.5, 0.99999994, -0.99999994]])
    else:
        mag = 10.0 ** rng.uniform(-6, 6, 1 << 20)
        x = np.concatenate([rng.standard_normal(1 << 20) * mag, [0.0, 0.0, 1.0, 0.0, 0.0, -1.0, -1.0, 0.0, 1.0, 1.0,
                                                                  -1.0, -1.0, 3.0, -3.0]])
    x = x.astype(np.float32)
    assert_bitexact(probe(pkg, ctx, op, x), O.math_array(op, x), f"device vs oracle, op {op}")


def test_bf16_conversion_rounds_to_nearest_even(pkg, ctx):
    """The learned network's float -> bf16 conversion (v_cvt_pk_bf16_f32)
    rounds to nearest, ties to even, as the oracle's bf16 rounding does: over
    random bit patterns of normal floats of both signs, exact ties with an even
    and an odd last kept bit, and values that round up to infinity.  Zero and
    infinity pass through, a NaN stays a NaN; a denormal gives its rounded
    value or a zero of its sign (no network value is that small)."""
    rng = np.random.default_rng(41)
    normal = rng.integers(0x00800000, 0x7F800000, size=1 << 20, dtype=np.uint32)
    sign = rng.integers(0, 2, size=normal.size, dtype=np.uint32) << 31
    ties = (rng.integers(0x0080, 0x7F80, size=4096, dtype=np.uint32) << 16) | 0x8000
    edge = np.array([0x7F7FFFFF, 0x7F7F8000, 0x7F7F7FFF, 0x3F808000, 0x3F818000, 0x00800000], np.uint32)
    bits = np.concatenate([normal | sign, ties, ties | 0x80000000, edge, edge | 0x80000000])
    x = bits.view(np.float32)
    u = bits.astype(np.uint64)
    want = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    assert_bitexact(probe(pkg, ctx, 7, x), want, "bf16 rounding")
    sp = np.array([0.0, -0.0, np.inf, -np.inf, 1.0, -2.0], np.float32)
    assert_bitexact(probe(pkg, ctx, 7, sp), sp, "bf16 of exactly representable values")
    assert np.isnan(probe(pkg, ctx, 7, np.array([np.nan], np.float32))).all()
    den = rng.integers(1, 0x00800000, size=4096, dtype=np.uint32)
    d = den.astype(np.uint64)
    dwant = (((d + 0x7FFF + ((d >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    got = probe(pkg, ctx, 7, den.view(np.float32))
    assert ((got.view(np.uint32) == dwant.view(np.uint32)) | (got == 0.0)).all()
