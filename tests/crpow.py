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


def cr_powf(x, y) -> np.float32:
    """Correctly rounded float32 x^y (x, y float32): the float32 nearest to the
    60-digit decimal value (ties to even), chosen among the neighbours of its
    double rounding."""
    x, y = np.float32(x), np.float32(y)
    if x == 0:
        return np.float32(0.0)
    v = (Decimal(float(y)) * Decimal(float(x)).ln()).exp()
    f = np.float32(float(v))
    cands = [np.nextafter(f, np.float32(0)), f, np.nextafter(f, np.float32(np.inf))]
    dist = [abs(Decimal(float(c)) - v) for c in cands]
    best = min(range(3), key=lambda i: (dist[i], int(cands[i].view(np.uint32)) & 1))
    return cands[best]


def sample_inputs_f32(n: int, seed: int = 0) -> np.ndarray:
    """Float hydraulic radii and wide-range floats, plus exact cubes and powers of 2."""
    rng = np.random.default_rng(seed)
    xs = [10 ** rng.uniform(-30, 22, n // 2), rng.uniform(0, 60, n // 2),
          np.array([1.0, 8.0, 27.0, 2.0 ** 30, 2.0 ** -30, 0.125, 1e-38, 1e22, 3.0, 64.0, 2.0 ** -149])]
    return np.concatenate(xs).astype(np.float32)
