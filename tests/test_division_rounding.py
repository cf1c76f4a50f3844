"""Ray set-up divides in float32 where the reference divides in float64 and rounds to float32
(core/ray.go:103-148).  The traversal kernels rely on the two giving identical bits for
float32 operands (double rounding is innocuous for division when 53 >= 2*24 + 2); this checks
that claim on random, extreme and special operands with numpy's IEEE float32 division."""
import numpy as np


def _operands(rng, n):
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    # random bit patterns cover every exponent, subnormals, infinities and NaNs
    c = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    d = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 1e-45, -1e-45, 1.1754944e-38, 3.4028235e38,
                        2.0**-149, 2.0**127, 1.0 + 2.0**-23, 1.0 - 2.0**-24, 3.0, 7.0, 0.1], np.float32)
    sa, sb = np.meshgrid(special, special)
    return np.concatenate([a, c, sa.ravel()]), np.concatenate([b, d, sb.ravel()])


def _same(x, y):
    xb, yb = x.view(np.uint32), y.view(np.uint32)
    both_nan = np.isnan(x) & np.isnan(y)
    return (xb == yb) | both_nan


def test_quotient_float64_then_float32_equals_float32():
    rng = np.random.default_rng(7)
    a, b = _operands(rng, 2_000_000)
    with np.errstate(all="ignore"):
        ref = (a.astype(np.float64) / b.astype(np.float64)).astype(np.float32)
        f32 = a / b
    assert _same(ref, f32).all()


def test_reciprocal_float64_then_float32_equals_float32():
    rng = np.random.default_rng(8)
    _, b = _operands(rng, 2_000_000)
    with np.errstate(all="ignore"):
        ref = (1.0 / b.astype(np.float64)).astype(np.float32)
        f32 = np.float32(1.0) / b
    assert _same(ref, f32).all()
