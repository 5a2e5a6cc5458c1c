"""Writes tests/golden/converged_cases.npz: inputs of the converged-rule
residual F(h) = Qmc(h) - Qmn(h) whose leftmost downcrossing is easy to get
wrong, with that downcrossing found by brute force.

Categories (cat):
  0  F(1 cm) <= 0 and a positive band of F narrower than a x1.25 depth grid:
     a scan of sample points (x1.25 from 1 cm, visiting bankfull) returns
     1 cm, the true leftmost downcrossing is the band's upper end;
  1  F(1 cm) <= 0 and F positive nowhere: the answer is 1 cm;
  2  the root lies above bankfull (floodplain), F > 0 at bankfull;
  3  an in-bank root and a second, floodplain root: the leftmost is in bank;
  4  F(1 cm) <= 0 with a band above bankfull.
The brute force is a dense scan (CASE_PTS points per decade on [1 cm, 300 m],
plus bankfull and the first point above it), the first sign change from > 0
to <= 0 refined by bisection to two ulps.  It uses only the oracle's point
residual (oracle/liborc.so, a plain restatement of the reference's formulas),
never its converged search, which is what the fixture pins.  Run from the
repository root:  python3 tests/golden/make_converged_cases.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from conftest import load_oracle  # noqa: E402

orc = load_oracle()
DT = 300.0
CASE_PTS = 20000       # points per decade of the brute force
SCREEN_PTS = 400


def bankfull(p):
    bw, tw, cs = p[1], p[2], p[6]
    z = 1.0 if cs == 0.0 else 1.0 / cs
    return bw / 1e-5 if bw > tw else (bw / (2 * z) if bw == tw else (tw - bw) / (2 * z))


def F(p, x, h):
    return orc.residual(p, *x, h, dt=DT)[0]


def grid(p, pts):
    bfd = bankfull(p)
    g = np.geomspace(0.01, 300.0, int(pts * np.log10(300.0 / 0.01)) + 1)
    extra = [bfd, np.nextafter(bfd, np.inf) * (1 + 1e-12)] if (p[3] > 0 and 0.01 < bfd < 300) else []
    return np.unique(np.concatenate([g, extra]))


def first_downcrossing(p, x, hs):
    f = np.array([F(p, x, h) for h in hs])
    if not (f > 0).any():
        return 0.01, f
    if f[0] <= 0:
        k0 = int(np.argmax(f > 0))
    else:
        k0 = 0
    rest = np.nonzero(f[k0:] <= 0)[0]
    if rest.size == 0:
        return None, f
    k = k0 + int(rest[0])
    a, b = hs[k - 1], hs[k]
    for _ in range(200):
        m = a + (b - a) / 2
        if not (a < m < b):
            break
        if F(p, x, m) > 0:
            a = m
        else:
            b = m
    return a + (b - a) / 2, f


def grid125_root(p, x):
    """The sample-point search a band can defeat: x1.25 from 1 cm, visiting
    bankfull, to the first F > 0, then to the first F <= 0, then bisection."""
    bfd = bankfull(p)
    h, used = 0.01, not (p[3] > 0 and bfd > 0.01)
    k = 0
    while not F(p, x, h) > 0:
        k += 1
        nh = h * 1.25
        if not used and nh >= bfd:
            nh, used = bfd, True
        if nh > 1e4 or k > 400:
            return 0.01
        h = nh
    a = b = h
    while True:
        nb = b * 1.25
        if not used and nb >= bfd:
            nb, used = bfd, True
        if not F(p, x, nb) > 0:
            b = nb
            break
        a = b = nb
    for _ in range(200):
        m = a + (b - a) / 2
        if not (a < m < b):
            break
        if F(p, x, m) > 0:
            a = m
        else:
            b = m
    return a + (b - a) / 2


def draw(rng):
    p = [rng.uniform(500, 5000), rng.uniform(1, 50), 0, 0, rng.uniform(0.03, 0.08), 0,
         rng.uniform(1, 3), float(np.clip(np.exp(rng.normal(np.log(1e-3), 1)), 1e-4, 1e-2))]
    p[2] = p[1] + rng.uniform(2, 20)
    p[3] = p[2] + rng.uniform(5, 100)
    p[5] = p[4] * rng.uniform(1.2, 2)
    p = np.array(p, np.float64)
    Q = np.exp(rng.uniform(np.log(0.05), np.log(500)))
    kind = rng.integers(3)
    if kind == 0:                         # rising: upstream inflow jumps above qup + qdp
        qup, qdp = Q * rng.uniform(0.3, 1.5), Q * rng.uniform(0.05, 1.0)
        quc = qup + qdp * rng.uniform(0.5, 4.0)
    elif kind == 1:                       # flood: large flows, overbank roots
        qup, quc, qdp = Q * 20 * rng.uniform(0.7, 1.3, 3)
    else:
        qup, quc, qdp = Q * rng.uniform(0.7, 1.3, 3)
    ql = float(rng.lognormal(np.log(0.1) - 0.5, 1.0)) * (1.0 if rng.random() < 0.9 else 5.0)
    return p, np.array([qup, quc, qdp, ql], np.float64)


def main():
    rng = np.random.default_rng(2024)
    want = {0: 6, 1: 6, 2: 8, 3: 8, 4: 2}
    got = {k: [] for k in want}
    tries = 0
    while any(len(got[k]) < want[k] for k in want) and tries < 200000:
        tries += 1
        p, x = draw(rng)
        bfd = bankfull(p)
        hs = grid(p, SCREEN_PTS)
        r, f = first_downcrossing(p, x, hs)
        if r is None:
            continue
        neg0 = f[0] <= 0
        downs = np.nonzero((f[:-1] > 0) & (f[1:] <= 0))[0]
        cat = None
        if neg0 and r == 0.01:
            cat = 1
        elif neg0 and r > bfd:
            cat = 4
        elif neg0:
            if grid125_root(p, x) == 0.01:
                cat = 0
        elif len(downs) >= 2 and r <= bfd < hs[downs[-1] + 1]:
            cat = 3
        elif r > bfd * (1 + 1e-9):
            cat = 2
        if cat is None or len(got[cat]) >= want[cat]:
            continue
        rf, _ = first_downcrossing(p, x, grid(p, CASE_PTS))   # the fine brute force
        if rf is None:
            continue
        got[cat].append((p, x, rf, cat))
        print(f"cat {cat}: {len(got[cat])}/{want[cat]} after {tries} draws, root {rf:.6g}, bankfull {bfd:.4g}",
              flush=True)
    cases = [c for k in want for c in got[k]]
    np.savez(os.path.join(ROOT, "tests", "golden", "converged_cases.npz"),
             params=np.array([c[0] for c in cases]), inputs=np.array([c[1] for c in cases]),
             root=np.array([c[2] for c in cases]), cat=np.array([c[3] for c in cases], np.int32),
             dt=DT, brute_force_points_per_decade=CASE_PTS)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
