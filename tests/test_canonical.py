"""The canonical root of the converged rules, on the CPU.

CONVERGED and LEARNED end every solve whose root is a downcrossing of the
residual in its canonical cell: on the grid of depths whose significand has
its low 28 bits zero, the neighbouring grid points g < g+ with Fc(g) > 0 >=
Fc(g+) next to the root estimate, and the secant point of that cell (cut at
the floodplain's jump when the cell holds bankfull).  Fc is the library's
residual form, whose every operation is IEEE, so its signs at grid points are
the same bits everywhere and any two solvers that land in the rounding band of
the root return the same depth and outflow.

What pins it here:
* the oracle's R^(2/3), R^(5/3) of Fc within two ulps of the exact powers
  (decimal arithmetic) and bit for bit the library's;
* Fc (oracle restatement) bit for bit the library's residual (host build of
  the device code), and within rounding of the reference-form residual;
* on every solve: the cell's ends are neighbouring grid points, Fc changes
  sign from > 0 to <= 0 across them, the root lies in the cell and within
  1e-12 of the reference-form residual's own downcrossing (bisection to two
  ulps inside the cell);
* the root does not depend on the estimate it starts from (anything in the
  rounding band, the cell's ends, a point a few cells away);
* the brute-force downcrossings of tests/golden/converged_cases.npz (roots at
  the floodplain's jump included) to 1e-12.
"""
import ctypes as C
import os
import subprocess
from decimal import Decimal

import numpy as np
import pytest

from conftest import ROOT, golden, load_oracle
from crpow import cr_pow, sample_inputs

orc = load_oracle()
SHIM = os.path.join(ROOT, "tests", "native", "libsolver_shim.so")
pD = C.POINTER(C.c_double)
SPACING_BITS = 28


@pytest.fixture(scope="module")
def shim():
    if not os.path.exists(SHIM):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "native")], check=True)
    L = C.CDLL(SHIM)
    L.shim_pow23_53_faithful.restype = None
    L.shim_pow23_53_faithful.argtypes = [pD, pD, C.c_int]
    L.shim_residual.restype = C.c_double
    L.shim_residual.argtypes = [pD, pD, C.c_double, C.c_double, pD]
    return L


def segc_row(p):
    dx, bw, tw, twcc, n, ncc, cs, s0 = map(float, p)
    z = 1.0 if cs == 0.0 else 1.0 / cs
    bfd = bw / 0.00001 if bw > tw else (bw / (2.0 * z) if bw == tw else (tw - bw) / (2.0 * z))
    return np.array([dx, bw, twcc, n, ncc, s0, z, bfd, np.sqrt(s0) / n, np.sqrt(1.0 + z * z), np.sqrt(s0)])


def bits(x):
    return int(np.float64(x).view(np.uint64))


def from_bits(b):
    return float(np.uint64(b).view(np.float64))


def is_grid(x):
    return bits(x) & ((1 << SPACING_BITS) - 1) == 0


def cases(n, seed):
    """Routing-like solves: channels with floodplains, flows +-30 % of a
    common scale over six decades, lateral inflow, the previous depth."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        p = [rng.uniform(500, 5000), rng.uniform(1, 50), 0, 0, rng.uniform(0.03, 0.08), 0,
             rng.uniform(1, 3), 10 ** rng.uniform(-4, -2)]
        p[2] = p[1] + rng.uniform(2, 20)
        p[3] = p[2] + rng.uniform(5, 100)
        p[5] = p[4] * rng.uniform(1.2, 2)
        Q = 10 ** rng.uniform(-2, 4)
        x = [Q * rng.uniform(0.7, 1.3), Q * rng.uniform(0.7, 1.3), Q * rng.uniform(0.7, 1.3), rng.uniform(0, 1)]
        out.append((p, x))
    return out


def ulps(a, b):
    return abs(bits(a) - bits(b))


def exact_pow(x, num, den):
    """x^(num/den) with the exact rational exponent (the fast powers compute
    x^(2/3) itself, where the correctly rounded pow of the reference's
    arithmetic takes the double nearest 2/3), correctly rounded."""
    if x == 0.0:
        return 0.0
    return float(((Decimal(num) / Decimal(den)) * Decimal(float(x)).ln()).exp())


def test_fast_powers_are_faithful_and_the_library_s(shim):
    xs = np.ascontiguousarray([x for x in sample_inputs(3000, 11) if 2.0 ** -600 <= x <= 2.0 ** 600])
    lib = np.zeros(2 * xs.size)
    shim.shim_pow23_53_faithful(xs.ctypes.data_as(pD), lib.ctypes.data_as(pD), xs.size)
    worst = 0
    for i, x in enumerate(xs):
        p23, p53 = orc.pow2353_fast(x, variant="cr")
        assert (p23, p53) == (lib[2 * i], lib[2 * i + 1]), x            # the library's bits
        worst = max(worst, ulps(p23, exact_pow(x, 2, 3)), ulps(p53, exact_pow(x, 5, 3)))
    assert worst <= 2, worst
    # outside the iteration's range: the correctly rounded powers
    for x in (0.0, 1e-310, 2.0 ** -1001, 2.0 ** 1001, 1e308):
        p23, p53 = orc.pow2353_fast(x, variant="cr")
        assert p23 == (cr_pow(x, 2.0 / 3.0) if x else 0.0)


def test_canon_residual_is_the_library_s_and_the_reference_form(shim):
    rng = np.random.default_rng(5)
    for p, x in cases(400, 21):
        seg = segc_row(p)
        xi = np.ascontiguousarray(x, np.float64)
        for h in 10 ** rng.uniform(-2, 1.5, 5):
            f, q = orc.canon_residual(p, *x, h, variant="cr")
            ql = C.c_double()
            fl = shim.shim_residual(seg.ctypes.data_as(pD), xi.ctypes.data_as(pD), 300.0, h, C.byref(ql))
            assert (bits(f), bits(q)) == (bits(fl), bits(ql.value)), (p, x, h)
            fr, qr = orc.residual(p, *x, h, variant="cr")
            # the two forms round differently: within a few ulps of the flows they combine
            scale = abs(x[0]) + abs(x[1]) + abs(x[2]) + abs(x[3]) + abs(qr) + abs(qr - fr)
            assert abs(f - fr) <= 1e-13 * scale, (p, x, h, f, fr)


def _bisect_reference_form(p, x, a, b):
    """The reference-form residual's downcrossing in [a, b] to two ulps."""
    fa = orc.residual(p, *x, a, variant="cr")[0]
    fb = orc.residual(p, *x, b, variant="cr")[0]
    if not (fa > 0.0 and not fb > 0.0):
        return None
    while ulps(a, b) > 2:
        m = 0.5 * (a + b)
        if orc.residual(p, *x, m, variant="cr")[0] > 0.0:
            a = m
        else:
            b = m
    return 0.5 * (a + b)


def test_canonical_cell_properties():
    n_checked = 0
    orc.noncanon_count("cr", reset=True)
    for p, x in cases(1500, 22):
        q, h, v = orc.solve_converged(p, *x, 1.0, variant="cr")
        if h <= 0.01:
            continue
        a = from_bits(bits(h) & ~((1 << SPACING_BITS) - 1))
        b = from_bits(bits(a) + (1 << SPACING_BITS))
        if h == b:                                   # the root on the upper grid point
            a, b = from_bits(bits(b) - (1 << SPACING_BITS)), b
        assert is_grid(a) and is_grid(b) and a <= h <= b
        fa, _ = orc.canon_residual(p, *x, a, variant="cr")
        fb, _ = orc.canon_residual(p, *x, b, variant="cr")
        bfd = segc_row(p)[7]
        if not (a <= bfd < b):                       # the jump cells are pinned by the golden cases
            assert fa > 0.0 and not fb > 0.0, (p, x, a, b, fa, fb)
            r = _bisect_reference_form(p, x, a, b)
            assert r is not None, (p, x, a, b)
            assert abs(h - r) <= 1e-12 * h, (p, x, h, r)
            n_checked += 1
    assert n_checked > 1000
    assert orc.noncanon_count("cr") == 0


def test_canonical_root_does_not_depend_on_the_estimate():
    for p, x in cases(500, 23):
        q, h, v = orc.solve_converged(p, *x, 1.0, variant="cr")
        if h <= 0.011:
            continue
        a = from_bits(bits(h) & ~((1 << SPACING_BITS) - 1))
        b = from_bits(bits(a) + (1 << SPACING_BITS))
        spacing = b - a
        starts = [h, h * (1 + 1e-15), h * (1 - 1e-15), h * (1 + 1e-13), h * (1 - 1e-13), a, b,
                  0.5 * (a + b), a + 2.5 * spacing, a - 2.5 * spacing]
        for r in starts:
            hr, qr = orc.canonical_root(p, *x, r, variant="cr")
            assert (bits(hr), bits(qr)) == (bits(h), bits(q)), (p, x, r, hr, h)


def test_golden_downcrossings_to_1e_12():
    z = golden("converged_cases.npz")
    dt = float(z["dt"])
    for i, (p, x, root, cat) in enumerate(zip(z["params"], z["inputs"], z["root"], z["cat"])):
        q, h, v = orc.solve_converged(p, *x, 1.0, dt=dt, variant="cr")
        assert abs(h - root) <= 1e-12 * root, (i, cat, h, root)
