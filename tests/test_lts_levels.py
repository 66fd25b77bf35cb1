"""The LTS level smoothing of the reference (solver-cpu.js:518-522: L in-place sweeps of
lv = min(lv, neighbour + 1)) equals the closed form the GPU kernel uses (min over tiles within
Manhattan distance L of raw level + distance) for raw levels in [0, L]; checked on random maps.
A dropped sweep or a wrong neighbour would make the two differ."""
import numpy as np
import pytest


def sweeps(lv0, L):
    ty, tx = lv0.shape
    lv = lv0.copy().reshape(-1)
    for _ in range(L):
        for t in range(tx * ty):
            x, y = t % tx, t // tx
            if x > 0: lv[t] = min(lv[t], lv[t - 1] + 1)
            if x < tx - 1: lv[t] = min(lv[t], lv[t + 1] + 1)
            if y > 0: lv[t] = min(lv[t], lv[t - tx] + 1)
            if y < ty - 1: lv[t] = min(lv[t], lv[t + tx] + 1)
    return lv.reshape(ty, tx)


def closure(lv0, L):
    ty, tx = lv0.shape
    out = lv0.copy()
    for y in range(ty):
        for x in range(tx):
            best = lv0[y, x]
            for dy in range(-L, L + 1):
                for dx in range(-(L - abs(dy)), L - abs(dy) + 1):
                    b, a = y + dy, x + dx
                    if 0 <= b < ty and 0 <= a < tx:
                        best = min(best, lv0[b, a] + abs(dx) + abs(dy))
            out[y, x] = best
    return out


@pytest.mark.parametrize("L,seed", [(1, 0), (2, 1), (3, 2), (4, 3), (4, 4)])
def test_closure_equals_sweeps(L, seed):
    rng = np.random.default_rng(seed)
    lv0 = rng.integers(0, L + 1, size=(13, 17))
    lv0[rng.random(lv0.shape) < 0.6] = L   # mostly coarse, as in practice
    a, b = sweeps(lv0, L), closure(lv0, L)
    np.testing.assert_array_equal(a, b)
    # the result is smooth: face neighbours differ by at most one level
    assert np.abs(np.diff(a, axis=0)).max() <= 1 and np.abs(np.diff(a, axis=1)).max() <= 1


def test_fewer_sweeps_differ():
    """One sweep short of L is not enough on a worst-case map (the check can fail)."""
    L = 4
    lv0 = np.full((1, 12), L)
    lv0[0, 0] = 0
    lv0[0, 11] = 0
    assert not np.array_equal(sweeps(lv0, L - 3), closure(lv0, L))
