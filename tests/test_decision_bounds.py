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


# ------------------------------------------------------------------ 2D crossing
# The 2D crossing-parity test (crossing_odd in dpm2d.hip) takes floor(x / L),
# round(x / L) and the sign of (djx - dix)(0 - diy) / (djy - diy) + dix from
# products with reciprocals where the result is certain, else from the
# correctly rounded division.  Here the fast forms, with the reciprocal's
# error (up to one ulp on the device) pushed to either end, against the
# divisions themselves, on values packed around every boundary.

F = np.float32


def fast_floor(x, L):
    q = F(x * F(F(1) / L))
    f = F(np.floor(q))
    m = F(F(1e-6) * F(F(1) + F(abs(q))))
    if F(q - f) > m and F(F(f + F(1)) - q) > m:
        return f
    return F(np.floor(F(x / L)))


def fast_round(x, L):
    q = F(x * F(F(1) / L))
    m = F(F(1e-6) * F(F(1) + F(abs(q))))
    h = F(F(abs(q)) - F(np.floor(F(abs(q)))))
    if abs(F(h - F(0.5))) > m:
        return F(np.sign(q) * np.floor(abs(q) + 0.5))
    r = F(x / L)
    return F(np.sign(r) * np.floor(abs(r) + F(0.5)))


def fast_sign(a, c, dix, rcp_err):
    rc = F(F(F(1) / c) * F(1 + rcp_err))
    qa = F(a * rc)
    ta = F(qa + dix)
    if abs(ta) > F(F(1e-6) * F(abs(qa) + abs(dix))):
        return bool(ta > 0)
    return bool(F(0) < F(F(a / c) + dix))


def c_round(v):
    return F(np.sign(v) * np.floor(abs(v) + F(0.5)))


def test_quotient_floor_and_round_equal_the_divisions():
    rng = np.random.default_rng(11)
    for L in F([9.969277, 38.4, 1.0, 153.6, 7.25]):
        # x near k L and near (k + 1/2) L, walked over floats
        for k in np.arange(-3, 4):
            for base in (k * L, (k + 0.5) * L):
                if base == 0:
                    continue  # (the bit walk would leave the floats; the samples below cover 0)
                c = np.array([base], F).view(np.int32)[0]
                xs = (np.arange(c - 300, c + 301).astype(np.int32)).view(F)
                for x in xs:
                    assert fast_floor(x, L) == F(np.floor(F(x / L))), (x, L)
                    assert fast_round(x, L) == c_round(F(x / L)), (x, L)
        for x in F(rng.uniform(-4 * L, 4 * L, 2000)):
            assert fast_floor(x, L) == F(np.floor(F(x / L)))
            assert fast_round(x, L) == c_round(F(x / L))


def test_crossing_sign_equals_the_division():
    rng = np.random.default_rng(12)
    eps = 2.0 ** -23
    n = 0
    for _ in range(20000):
        a = F(rng.uniform(-3, 3))
        c = F(rng.uniform(0.01, 3) * rng.choice([-1, 1]))
        q = F(a / c)
        # dix close to -q, so that tt sits near 0, and some far away
        dix = F(-q + F(rng.normal() * abs(q) * 10.0 ** rng.integers(-9, 0)))
        exact = bool(F(0) < F(q + dix))
        for err in (-eps, 0.0, eps):
            assert fast_sign(a, c, dix, err) == exact, (a, c, dix, err)
        n += 1
    # exact ties and zeros
    for a, c, dix in ((F(0), F(1), F(0)), (F(1), F(1), F(-1)), (F(3), F(-2), F(1.5)), (F(1e-30), F(1e-8), F(0))):
        exact = bool(F(0) < F(F(a / c) + dix))
        for err in (-eps, 0.0, eps):
            assert fast_sign(a, c, dix, err) == exact
