"""The step-size root of the integrator core (csrc/bdf_core.h root_n): x^(1/n)
from an exact split of the binary exponent, x = m 2^e with e = k n + r
(C division, truncating toward zero), and FP32 log2/exp2 of m 2^r only.

CVODE forms its step ratios as (BIAS * dsm)^(1/L) (cvode.c, "Mathematical
considerations" of the CVODE guide); the kernel's approximation must agree
with the exact root to FP32 accuracy over the whole double range, for every
order 1..6 (L <= QMAX + 1), including negative exponents where truncating
division gives a negative remainder (a dropped sign there would be off by
a factor 2^(2r/n))."""
import math

import numpy as np
import pytest


def root_n(x: float, n: int) -> float:
    if x == 0.0:
        return 0.0
    m, e = math.frexp(x)                      # x = m 2^e, m in [0.5, 1)
    k = int(e / n)                            # C: truncates toward zero
    r = e - k * n
    f = (np.float32(r) + np.log2(np.float32(m))) * (np.float32(1.0) / np.float32(n))
    return math.ldexp(float(np.exp2(np.float32(f))), k)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_root_matches_exact(n):
    rng = np.random.default_rng(n)
    xs = 10.0 ** rng.uniform(-300, 300, 2000)
    # (within an FP32 rounding of DBL_MAX the n = 1 root may round up to inf;
    # step-control inputs never come near it)
    xs = np.concatenate([xs, [5e-324, 2.2250738585072014e-308, 1.0, 0.5, 2.0, 1e307]])
    for x in xs:
        got = root_n(float(x), n)
        ref = math.exp(math.log(x) / n)
        assert abs(got - ref) <= 4e-7 * ref, (x, n, got, ref)


def test_root_edge_values():
    assert root_n(0.0, 3) == 0.0
    assert root_n(1.0, 5) == pytest.approx(1.0, rel=1e-7)
    assert root_n(8.0, 3) == pytest.approx(2.0, rel=1e-7)
    assert root_n(2.0 ** -99, 3) == pytest.approx(2.0 ** -33, rel=1e-7)
