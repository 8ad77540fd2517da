"""The table-driven FP64 exp and log of the device code (csrc/nsw6_math.h), run on the
host through the test library: tables recomputed to 80 digits, accuracy against
high-precision decimal values and the C library over the whole double range, and the
C library's special values."""
from decimal import Decimal, getcontext

import numpy as np

from oracle import oracle as O


def _run(which, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    assert O.grad_lib().grad_math_tang(which, x.ctypes.data, y.ctypes.data, x.size) == 0
    return y


def _ulps(y, ref):
    """|y - ref| in units of the last place of ref (ref finite, nonzero)."""
    return np.abs(y - ref) / np.spacing(np.abs(ref))


def test_tables_are_correctly_rounded():
    e = np.empty(64)
    lt = np.empty(130)
    O.grad_lib().grad_math_tables(e.ctypes.data, lt.ctypes.data)
    getcontext().prec = 80
    for j in range(64):
        assert e[j] == float(Decimal(2) ** (Decimal(j) / 64)), j
    for j in range(65):
        exact = (Decimal(64 + j) / 128).ln()
        hi = float(exact)
        assert lt[2 * j] == hi, j
        assert lt[2 * j + 1] == float(exact - Decimal(hi)), j


def test_exp_against_decimal():
    getcontext().prec = 60
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-700, 700, 1500), rng.uniform(-1, 1, 500), rng.uniform(-1e-6, 1e-6, 200),
                        -np.arange(0, 700, 7.3), np.log(2) * np.arange(-20, 21) / 64])
    y = _run(0, x)
    ref = np.array([float(Decimal(float(v)).exp()) for v in x])
    u = _ulps(y, ref)
    assert u.max() <= 1.5, (x[u.argmax()], u.max())


def test_log_against_decimal():
    getcontext().prec = 60
    rng = np.random.default_rng(8)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 1500), rng.uniform(0.5, 2.0, 500),
                        1.0 + rng.uniform(-1e-3, 1e-3, 200), 1.0 + 2.0 ** -np.arange(1, 50),
                        1.0 - 2.0 ** -np.arange(2, 50), 0.5 + np.arange(65) / 128])
    x = x[x != 1.0]
    y = _run(1, x)
    ref = np.array([float(Decimal(float(v)).ln()) for v in x])
    u = _ulps(y, ref)
    assert u.max() <= 1.5, (x[u.argmax()], u.max())
    assert _run(1, np.array([1.0]))[0] == 0.0


def test_bulk_against_the_c_library():
    rng = np.random.default_rng(9)
    x = rng.uniform(-745.0, 709.7, 2_000_000)
    u = _ulps(_run(0, x), np.exp(x))
    assert u.max() <= 2.0
    x = 10.0 ** rng.uniform(-307.5, 308.2, 2_000_000)
    u = _ulps(_run(1, x), np.log(x))
    assert u.max() <= 2.0
    # pow(x, y) as the rate code forms it, exp(y log x), against the library's pow: the
    # log's error and the rounding of the product y log x carry |y log x| ulp of 1
    b = 10.0 ** rng.uniform(-12, 3, 200_000)
    p = rng.uniform(-2.5, 2.5, 200_000)
    rel = np.abs(_run(0, p * _run(1, b)) / np.power(b, p) - 1.0)
    assert np.all(rel <= (2.0 * np.abs(p * np.log(b)) + 3.0) * 2.3e-16)


def test_subnormal_and_extreme_arguments():
    tiny = np.array([5e-324, 1e-310, 2.2250738585072014e-308, 1.7976931348623157e308])
    u = _ulps(_run(1, tiny), np.log(tiny))
    assert u.max() <= 2.0
    # results near the bottom of the range lose bits to the subnormals as the library does
    x = np.array([-708.0, -709.5, -720.0, -740.0, -744.0, 709.0, 709.78])
    y, ref = _run(0, x), np.exp(x)
    assert np.all(np.abs(y - ref) <= 2.0 * np.spacing(np.abs(ref)))


def test_special_values():
    inf, nan = np.inf, np.nan
    e = _run(0, np.array([-inf, inf, nan, 0.0, -0.0, 710.0, 1e10, -746.0, -1e10, 1100.0, -1100.0]))
    assert e[0] == 0.0 and e[1] == inf and np.isnan(e[2])
    assert e[3] == 1.0 and e[4] == 1.0
    assert e[5] == inf and e[6] == inf and e[9] == inf
    assert e[7] == 0.0 and e[8] == 0.0 and e[10] == 0.0
    lg = _run(1, np.array([0.0, -0.0, inf, -1.0, -inf, nan, -5e-324]))
    assert lg[0] == -inf and lg[1] == -inf and lg[2] == inf
    assert np.isnan(lg[3]) and np.isnan(lg[4]) and np.isnan(lg[5]) and np.isnan(lg[6])

