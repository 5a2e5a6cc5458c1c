"""Correctly rounded x^y by exact decimal arithmetic (the judge of both pow
implementations): Decimal ln/exp are correctly rounded at the context
precision, 60 digits here, and float(Decimal) rounds to nearest-even."""
from decimal import Decimal, getcontext

import numpy as np

getcontext().prec = 60


def cr_pow(x: float, y: float) -> float:
    if x == 0.0:
        return 0.0
    return float((Decimal(y) * Decimal(x).ln()).exp())


def sample_inputs(n: int, seed: int = 0) -> np.ndarray:
    """Hydraulic radii and wide-range values, plus exact cubes and powers of 2."""
    rng = np.random.default_rng(seed)
    xs = [10 ** rng.uniform(-12, 12, n // 2), rng.uniform(0, 60, n // 2),
          np.array([1.0, 8.0, 27.0, 2.0 ** 30, 2.0 ** -30, 0.125, 1e-300, 1e300, 3.0, 64.0])]
    return np.concatenate(xs)
