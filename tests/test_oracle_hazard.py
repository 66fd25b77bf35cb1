"""The oracle's hazard products pinned to the reference's known answers (tests/run-hazard-tests.mjs:9-38):
ARR 2019 H1-H6 and FD2320 classes at and around their limits, force per width, and the tracker's
rules (class from simultaneous depth and speed, duration and arrival, films under 1 cm ignored,
thin fast water classed by depth alone)."""
import math

import numpy as np
import pytest

import oracle

VCUT = 0.05


@pytest.mark.parametrize("d,v,c", [(0.2, 1.0, 1), (0.3, 1.0, 1), (0.31, 0.5, 2), (0.2, 2.0, 2), (0.45, 1.2, 2),
                                   (1.0, 0.5, 3), (0.6, 1.2, 4), (1.5, 0.5, 4), (0.4, 2.1, 5), (2.5, 1.0, 5),
                                   (1.0, 3.9, 5), (4.5, 0.2, 6), (1.0, 4.5, 6), (2.0, 2.1, 6)])
def test_arr_classes(d, v, c):
    assert oracle.lib().or_hz_arr(d, v) == c


def test_fd2320_classes():
    L = oracle.lib()
    D = lambda d, v: L.or_hz_hr(d * (v + 0.5))
    U = lambda d, v: L.or_hz_hr(d * (v + 0.5) + (1 if d > 0.25 else 0))
    assert (D(0.5, 0.9), D(0.5, 1.1), D(0.5, 2.0), D(1.0, 1.6)) == (1, 2, 3, 4)
    assert U(0.2, 1) == D(0.2, 1) and U(0.3, 0.5) == 3 and D(0.3, 0.5) == 1


def test_force_per_width():
    L = oracle.lib()
    assert abs(L.or_hz_force(0.5, 2, 9.81) - (0.5 * 4 + 0.5 * 9.81 * 0.25)) < 1e-9
    assert abs(L.or_hz_force(1, 0, 9.81) - 4.905) < 1e-9


class _Tracker:
    """HazardTracker.update (hazard.js:77-80) over a one-cell oracle."""

    def __init__(self):
        self.s = oracle.CPUSolver(nx=1, ny=1, dx=1.0, z=np.zeros(1))
        self.t = None

    def update(self, t, d, v):
        dt = 0.0 if self.t is None else max(0.0, t - self.t)
        self.t = t
        oracle.lib().or_hazard_cell(self.s._h, 0, d, v if d >= VCUT else 0.0, t, dt)

    def read(self):
        return {k: v[0] for k, v in self.s.readHazard().items()}


def test_tracker_same_moment_duration_arrival():
    tr = _Tracker()
    for t, d, v in ((0, 0, 0), (60, 1.0, 0.2), (120, 0.1, 1.5), (180, 0.05, 0.1)):
        tr.update(t, d, v)
    h = tr.read()
    assert h["arr"] == 3 and oracle.lib().or_hz_arr(1.0, 1.5) == 5
    assert abs(h["dv"] - 0.2) < 1e-6
    assert h["duration"] == 60 and h["arrival"] == 60


def test_tracker_thin_films():
    tr = _Tracker()
    tr.update(0, 0.005, 3)
    tr.update(10, 0.005, 3)
    h = tr.read()
    assert h["arr"] == 0 and h["force"] == 0 and math.isnan(h["arrival"])
    tr = _Tracker()
    tr.update(0, 0.03, 5.5)
    h = tr.read()
    assert h["arr"] == 1 and h["dv"] == 0
