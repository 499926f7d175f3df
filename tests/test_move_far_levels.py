"""The level order of the far-jump move (CPU).

When some particle jumps R > 1 cells, the move runs the colour phases' slot
loops cell by cell in one workgroup.  In each colour phase the cells (colour coordinates
t) go in levels L(t) = sum_d m_d t_d, m_0 = 1, m_d = 1 + sum_{e<d} m_e q_e,
q_e = min(floor(2R/3), ncol_e - 1), levels 0 .. sum_d m_d (ncol_d - 1), the
t_0 of a level found from the other coordinates.  Two cells interact when
their colour coordinates differ by at most floor(2R/3) everywhere (their
cells are within 2R of each other).  Checked here by brute force: every cell
gets exactly one level, and any two interacting cells get different levels
in the reference's linear order (t_0 fastest).
"""
import itertools

import numpy as np
import pytest


def levels(ncol, R):
    D = len(ncol)
    q = (2 * R) // 3
    m = []
    for d in range(D):
        m.append(1 + sum(m[e] * min(q, ncol[e] - 1) for e in range(d)))
    lmax = sum(m[d] * (ncol[d] - 1) for d in range(D))
    outer = int(np.prod(ncol[1:]))
    out = {}
    for lv in range(lmax + 1):
        for i in range(outer):  # the kernel's enumeration, one thread per i
            t, rest, l0 = [0] * D, i, lv
            for d in range(1, D):
                t[d] = rest % ncol[d]
                rest //= ncol[d]
                l0 -= m[d] * t[d]
            t[0] = l0
            if 0 <= t[0] < ncol[0]:
                assert tuple(t) not in out
                out[tuple(t)] = lv
    return out, q


@pytest.mark.parametrize("ncol,R", [((11, 8), 2), ((14, 12), 4), ((6, 22), 3), ((7, 7), 20),
                                    ((1, 5), 2), ((3, 3, 3), 2), ((3, 3, 3), 5), ((5, 4, 3), 3),
                                    ((4, 1, 2), 7), ((2, 6, 2), 9)])
def test_far_levels_keep_the_reference_order(ncol, R):
    lv, q = levels(ncol, R)
    assert len(lv) == int(np.prod(ncol))
    lin = {t: int(np.ravel_multi_index(t, ncol, order="F")) for t in lv}
    for a, b in itertools.combinations(lv, 2):
        if max(abs(x - y) for x, y in zip(a, b)) <= q:
            assert lv[a] != lv[b]
            assert (lv[a] < lv[b]) == (lin[a] < lin[b])
