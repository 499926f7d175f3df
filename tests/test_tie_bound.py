"""CPU check of the bound the injection's nearest-neighbour search relies on.

The GPU search compares squared distances s and reproduces the reference's
argmin of the rounded distance sqrt(s) (first in order on ties).  It assumes
that an s above best * (1 + 2^-50) never rounds to the same sqrt as best, so
only squares within that range need the exact sqrt comparison.  Here every
double above best is walked ulp by ulp until its sqrt differs, for many best
values (random, powers of two, values next to powers of two, subnormal-free
extremes), and the last tying square must stay within the bound.
"""
import numpy as np

LIM = 1.0 + 2.0 ** -50


def last_tie(best):
    m = np.sqrt(best)
    s = best
    while True:
        n = np.nextafter(s, np.inf)
        if np.sqrt(n) != m:
            return s
        s = n


def test_ties_stay_within_bound():
    rng = np.random.default_rng(0)
    vals = list(rng.uniform(1e-12, 1e3, 400))
    vals += [2.0 ** k for k in range(-40, 20)]
    vals += [np.nextafter(2.0 ** k, 0.0) for k in range(-40, 20)]
    vals += [np.nextafter(2.0 ** k, np.inf) for k in range(-40, 20)]
    vals += [1e-300, 1e300, 3.0, 0.5 * (1 + 2.0 ** -52)]
    for best in vals:
        best = np.float64(best)
        s = last_tie(best)
        assert s <= best * np.float64(LIM), (best, s)
        # the first square above the bound already rounds differently
        above = np.nextafter(best * np.float64(LIM), np.inf)
        assert np.sqrt(above) > np.sqrt(best)


def test_argmin_on_squares_matches_rounded_distances():
    """Brute force: argmin over sqrt with first-occurrence ties equals the
    squared-distance argmin plus the near-tie pass, on crafted near ties."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        base = rng.uniform(1e-6, 1.0)
        n = int(rng.integers(2, 12))
        s = np.array([np.float64(base)] * n)
        for i in range(n):  # perturb by a few ulps either side
            for _ in range(int(rng.integers(0, 4))):
                s[i] = np.nextafter(s[i], np.inf if rng.random() < 0.5 else 0.0)
        rng.shuffle(s)
        d = np.sqrt(s)
        want = int(np.argmin(d))  # numpy argmin: first occurrence of the minimum
        best = s.min()
        got = int(np.nonzero(s == best)[0][0])
        m = np.sqrt(best)
        near = [i for i in range(n) if best < s[i] <= best * LIM and np.sqrt(s[i]) == m]
        if near:
            got = min(got, min(near))
        assert got == want
