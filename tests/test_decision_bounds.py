"""The square-root-free form of the 3D adhesion decision (AdhCut in
dpm3d.hip) against the decision it replaces, on the CPU.

The reference accepts a vertex pair when sqrt(d2) < 2 l0 and
sqrt(d2) > 1e-12 (Cell3D_Kernel.cl:348-353), with a correctly rounded float
square root.  The float steps decide the same thing as d2 < m^2 and
d2 > m'^2 in double, m (m') the midpoint between the float threshold and the
float below (above) it.  Here both forms are evaluated on every float d2 in a
window of a few thousand floats around each boundary, for thresholds across
the range of rest lengths and for the 1e-12 floor; numpy's float32 square
root is the IEEE correctly rounded one.
"""
import numpy as np


def below(c):
    return (np.array([c], np.float32).view(np.int32) - 1).view(np.float32)[0]


def above(c):
    return (np.array([c], np.float32).view(np.int32) + 1).view(np.float32)[0]


def bound_lt(c):
    """d2 < bound_lt(c) <=> RN(sqrt(d2)) < c (the AdhCut upper bound)."""
    m = 0.5 * (np.float64(below(c)) + np.float64(c))
    return m * m


def bound_gt(c):
    """d2 > bound_gt(c) <=> RN(sqrt(d2)) > c (the AdhCut lower bound)."""
    m = 0.5 * (np.float64(c) + np.float64(above(c)))
    return m * m


def window(center, n=4096):
    """The n floats below and above the float nearest `center`."""
    c = np.array([center], np.float32).view(np.int32)[0]
    return (np.arange(c - n, c + n + 1, dtype=np.int64).astype(np.int32)).view(np.float32)


def thresholds():
    rng = np.random.default_rng(7)
    # twice rest lengths over the range the tissues use, powers of two (the
    # float below is in the lower binade), and exact squares
    cs = list((2.0 * rng.uniform(1e-3, 3.0, 400)).astype(np.float32))
    cs += [np.float32(2.0 ** k) for k in range(-12, 6)]
    cs += [np.float32(k) for k in (1, 3, 9, 25, 0.25, 0.5625)]
    return cs


def test_upper_bound_equals_the_sqrt_decision():
    for c in thresholds():
        b = bound_lt(c)
        x = window(b)
        direct = np.sqrt(x) < c
        cut = x.astype(np.float64) < b
        assert np.array_equal(direct, cut), c
        assert direct.any() and not direct.all(), c  # the window spans the boundary


def test_lower_bound_equals_the_sqrt_decision():
    tiny = np.float32(1e-12)
    b = bound_gt(tiny)
    x = window(b)
    direct = np.sqrt(x) > tiny
    cut = x.astype(np.float64) > b
    assert np.array_equal(direct, cut)
    assert direct.any() and not direct.all()
    # and the whole pair decision on zero, subnormal and huge squares
    c = np.float32(0.62)
    for x in np.array([0.0, 1e-45, 1e-40, 1e-24, 1.0000001e-24, 0.3844, 0.38440004, 1e30], np.float32):
        s = np.sqrt(x)
        direct = bool(s < c and s > tiny)
        cut = bool(np.float64(x) < bound_lt(c) and np.float64(x) > b)
        assert direct == cut, x


def test_no_float_square_lands_on_a_midpoint():
    # the boundary m^2 is never a float, so no rounding tie can arise
    for c in thresholds():
        b = bound_lt(c)
        assert np.float64(np.float32(b)) != b, c
