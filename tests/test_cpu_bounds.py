"""CPU checks of the rounding bound that the molecule-per-lane tile kernel adds
to a frame's guard band when it evaluates distances in the expanded form
relative to the chunk centre c:

    |y' - s'|^2 = (|s'|^2 - 2 y'.s') + |y'|^2,   y' = fl(y - c), s' = fl(s - c),

with |s'|^2 staged beside s', three FMAs per pair against the lane's -2 y',
and |y'|^2 added once.  The claimed bound on the distance between that FP32
value and the exact d^2 of the same FP32 coordinates is

    u (Y + S) (2 d + 7 (Y + S)),   u = 2^-24, Y = |y - c|, S = |s - c|.

The kernel's FP32 arithmetic is emulated with numpy: float32 operations round
once; an FMA is the exact float64 product plus the addend, rounded to
float32 (the float64 sum adds at most 2^-53 relative before that rounding).
"""
import numpy as np

U = 2.0 ** -24


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def fma32(a, b, c):
    return f32(a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64))


def expanded_d2(y, s, c):
    """The kernel's sequence for one atom y of the lane and one staged atom s
    (arrays of shape (n, 3), float32)."""
    sp = f32(s - c)                      # staging: s' = s - c
    w = fma32(sp[:, 0], sp[:, 0], fma32(sp[:, 1], sp[:, 1], f32(sp[:, 2] * sp[:, 2])))
    ym = f32(np.float32(-2.0) * f32(y - c))   # the lane keeps -2 y'
    t = fma32(ym[:, 0], sp[:, 0], fma32(ym[:, 1], sp[:, 1], fma32(ym[:, 2], sp[:, 2], w)))
    yy2 = f32(np.float32(0.25) * fma32(ym[:, 0], ym[:, 0], fma32(ym[:, 1], ym[:, 1], f32(ym[:, 2] * ym[:, 2]))))
    return f32(t + yy2)


def adversarial(rng, n, box=110.0, yr=6.0, dmax=10.5):
    """Chunk centres anywhere in a 110 A box, lane atoms up to 6 A from the
    centre (a chunk's half-diagonal), staged atoms 0-10.5 A from the lane
    atom, all in float32 as the kernel holds them."""
    c = f32(rng.uniform(0.0, box, (n, 3)))
    y = f32(c + rng.uniform(-yr, yr, (n, 3)))
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    d = rng.uniform(0.0, dmax, n)
    s = f32(y + d[:, None] * v)
    return y, s, c


def errors(y, s, c):
    d2 = np.sum((y.astype(np.float64) - s.astype(np.float64)) ** 2, axis=1)   # exact to 2^-52
    v = expanded_d2(y, s, c).astype(np.float64)
    Y = np.linalg.norm(y.astype(np.float64) - c, axis=1)
    S = np.linalg.norm(s.astype(np.float64) - c, axis=1)
    return np.abs(v - d2), np.sqrt(d2), Y, S


def test_expanded_form_within_bound():
    rng = np.random.default_rng(2026)
    y, s, c = adversarial(rng, 400_000)
    err, d, Y, S = errors(y, s, c)
    bound = U * (Y + S) * (2.0 * d + 7.0 * (Y + S)) + 1e-12
    ratio = err / bound
    assert ratio.max() <= 1.0, ratio.max()
    # not vacuous: the largest errors come within a small factor of the bound
    assert ratio.max() > 0.05, ratio.max()


def test_expanded_form_needs_the_quadratic_term():
    """The (Y + S)^2 part (norms, FMA chain, last sum) is not covered by the
    coordinate term alone: dropping it fails on these inputs."""
    rng = np.random.default_rng(2027)
    y, s, c = adversarial(rng, 400_000)
    err, d, Y, S = errors(y, s, c)
    assert (err > U * (Y + S) * 2.0 * d + 1e-12).any()


def test_direct_form_error_is_smaller_near_contact():
    """Sanity of the emulation: for the direct form (y - s)^2 on the same
    float32 coordinates the error is at most about 4 u d^2 (no cancellation)."""
    rng = np.random.default_rng(2028)
    y, s, c = adversarial(rng, 100_000)
    dx = f32(y - s)
    v = fma32(dx[:, 0], dx[:, 0], fma32(dx[:, 1], dx[:, 1], f32(dx[:, 2] * dx[:, 2]))).astype(np.float64)
    d2 = np.sum((y.astype(np.float64) - s.astype(np.float64)) ** 2, axis=1)
    assert (np.abs(v - d2) <= 5.0 * U * d2 + 1e-12).all()
