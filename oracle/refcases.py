"""The reference's test scenarios and closed forms, restated -- TEST INFRASTRUCTURE ONLY.

  rng, cloudburst, dambreak, thacker   src/scenarios.js:4-84
  damBreak (Stoker / Ritter)           src/analytic.js:5-27
  thackerPlanar                        src/analytic.js:30-40
  barrier_case                         the flood-barrier case of tests/gpu-suite.js:242-253
  gi_case, ga_case                     the GI and Green-Ampt cases of tests/gpu-suite.js:188-193, 332-345

They rebuild the inputs of the recorded fixtures (tests/results-cpu.json) so the
oracle can be checked against them. Loops follow the JS order where a random
stream is consumed; Math.round (half up) is floor(x + 0.5).
"""
import math

import numpy as np

MMHR = 1 / 1000 / 3600
G = 9.81


def rng(seed):
    """LCG of src/scenarios.js:4."""
    s = [seed & 0xFFFFFFFF]

    def nxt():
        s[0] = (s[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return s[0] / 4294967296

    return nxt


def _jsround(x):
    return math.floor(x + 0.5)


def cloudburst(n):
    """320 m urban catchment (src/scenarios.js:8-49)."""
    L = 320
    nx = ny = n
    dx = L / n
    rnd = rng(11)
    sw = 10
    streets = [40, 120, 200, 280]
    creekY = lambda x: 160 + 34 * math.sin(x / 58 + 0.6)
    edges = [0] + streets + [L]
    rects = []
    park = None
    for bi in range(len(edges) - 1):
        for bj in range(len(edges) - 1):
            x0 = edges[bi] + (sw / 2 + 3 if bi else 3)
            x1 = edges[bi + 1] - (sw / 2 + 3 if bi + 1 < len(edges) - 1 else 3)
            y0 = edges[bj] + (sw / 2 + 3 if bj else 3)
            y1 = edges[bj + 1] - (sw / 2 + 3 if bj + 1 < len(edges) - 1 else 3)
            if bi == 1 and bj == 3:
                park = (x0, x1, y0, y1)
                continue
            mx, my = (x0 + x1) / 2, (y0 + y1) / 2
            for lx0, lx1 in ((x0, mx - 2), (mx + 2, x1)):
                for ly0, ly1 in ((y0, my - 2), (my + 2, y1)):
                    cx, cy = (lx0 + lx1) / 2, (ly0 + ly1) / 2
                    if abs(cy - creekY(cx)) < 22:
                        continue
                    ix = 1 + rnd() * 4
                    iy = 1 + rnd() * 4
                    rects.append((lx0 + ix, lx1 - ix * 0.6, ly0 + iy, ly1 - iy * 0.6, 4 + _jsround(rnd() * 4) * 2.5))
    N = nx * ny
    z = np.zeros(N)
    mask = np.ones(N, dtype=np.uint8)
    man = np.zeros(N)
    inf = np.zeros(N)
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            x, y = (i + 0.5) * dx, (j + 0.5) * dx
            smooth = 12 - 0.012 * x + 0.000035 * (y - 160) ** 2
            zz = smooth + 0.5 * math.sin(x / 37 + 1.3) * math.sin(y / 29 + 0.4) + 0.25 * math.sin(x / 13 + 2) * math.cos(y / 17)
            nn, f = 0.035, 12 * MMHR
            if park and park[0] < x < park[1] and park[2] < y < park[3]:
                nn, f = 0.05, 25 * MMHR
            if any(abs(x - s) < sw / 2 or abs(y - s) < sw / 2 for s in streets):
                zz -= 0.15
                nn, f = 0.016, 0
            yc = creekY(x)
            d = abs(y - yc)
            if d < 7:
                bed = 12 - 0.012 * x + 0.000035 * (yc - 160) ** 2 - 1.6
                zz = min(zz, bed + (zz - bed) * (d / 7) ** 2)
                nn, f = 0.03, 4 * MMHR
            for r in rects:
                if r[0] < x < r[1] and r[2] < y < r[3]:
                    mask[k] = 0
                    break
            z[k], man[k], inf[k] = zz, nn, f
    return dict(nx=nx, ny=ny, dx=dx, z=z, mask=mask, manning=man, infil=inf, bc={"east": "open"},
                depth=lambda X, Y, Z: 0 * X)


def dambreak(n):
    """Valley dam break (src/scenarios.js:51-73)."""
    ny, nx = n, 2 * n
    Lx = 480
    dx = Lx / nx
    N = nx * ny
    z = np.zeros(N)
    mask = np.ones(N, dtype=np.uint8)
    man = np.full(N, 0.03)
    rnd = rng(5)
    rects = []
    bx = 236
    while bx < 400:
        by = 30
        while by < 214:
            if abs(by + 5 - 120) < 12:
                by += 22
                continue
            if rnd() < 0.18:
                by += 22
                continue
            a = bx + 12 + rnd() * 4
            b = by + 9 + rnd() * 3
            c = 4 + _jsround(rnd() * 3) * 2.5
            rects.append((bx, a, by, b, c))
            by += 22
        bx += 26
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            x, y = (i + 0.5) * dx, (j + 0.5) * dx
            z[k] = 6 - 0.006 * x + 0.00028 * (y - 120) ** 2 + 0.15 * math.sin(x / 23) * math.sin(y / 19)
            if abs(y - 120) < 6 and x > 200:
                man[k] = 0.016
            for r in rects:
                if r[0] < x < r[1] and r[2] < y < r[3]:
                    mask[k] = 0
                    break
    return dict(nx=nx, ny=ny, dx=dx, z=z, mask=mask, manning=man, infil=None, bc={"east": "open"},
                depth=lambda X, Y, Z: np.where(X < 70, 13.5 - Z, 0.0))


def damBreak(hL, hR, x0, g=G):
    """Exact Stoker (hR > 0) / Ritter (hR = 0) dam break h(x, t) (src/analytic.js:5-27)."""
    cL = math.sqrt(g * hL)
    hm = um = S = 0.0
    if hR > 0:
        f = lambda h: 2 * (cL - math.sqrt(g * h)) - (h - hR) * math.sqrt(g * (h + hR) / (2 * h * hR))
        lo, hi = hR, hL
        for _ in range(200):
            m = 0.5 * (lo + hi)
            if f(m) > 0:
                lo = m
            else:
                hi = m
        hm = 0.5 * (lo + hi)
        um = 2 * (cL - math.sqrt(g * hm))
        S = hm * um / (hm - hR)

    def h(x, t):
        if t <= 0:
            return hL if x < x0 else hR
        xi = (x - x0) / t
        if xi <= -cL:
            return hL
        if hR <= 0:
            if xi < 2 * cL:
                q = 2 * cL - xi
                return q * q / (9 * g)
            return 0.0
        cm = math.sqrt(g * hm)
        if xi < um - cm:
            q = 2 * cL - xi
            return q * q / (9 * g)
        if xi < S:
            return hm
        return hR

    return h


class ThackerPlanar:
    """Planar surface in a paraboloid (src/analytic.js:30-40)."""

    def __init__(self, L=4.0, a=1.0, h0=0.1, eta=0.5, g=G):
        self.L, self.a, self.h0, self.eta, self.g = L, a, h0, eta, g
        self.omega = math.sqrt(2 * g * h0) / a
        self.period = 2 * math.pi / self.omega

    def bed(self, x, y):
        X, Y = x - self.L / 2, y - self.L / 2
        return -self.h0 * (1 - (X * X + Y * Y) / (self.a * self.a))

    def depth(self, x, y, t):
        X, Y = x - self.L / 2, y - self.L / 2
        w = self.omega
        s = (self.eta * self.h0 / (self.a * self.a)) * (2 * X * np.cos(w * t) + 2 * Y * np.sin(w * t) - self.eta)
        return np.maximum(0, s - self.bed(x, y))

    def vel(self, t):
        w = self.omega
        return (-self.eta * w * math.sin(w * t), self.eta * w * math.cos(w * t))


def thacker_bed(nx, T=None):
    T = T or ThackerPlanar()
    dx = T.L / nx
    x = (np.arange(nx) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    return T.bed(X, Y).reshape(-1), dx, T


def barrier_case():
    """gpu-suite.js:242-253: a wall of barrier faces across a sloping 64 x 48 basin (crest 5.4 m, C 1.6)
    and a gap-free diagonal leg (crest 4.9 m, C 1.44), a dam-break start, the leg raised 0.5 m at 30 s."""
    nx, ny, dx = 64, 48, 2
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = (5 - 0.02 * I + 0.01 * np.sin(J / 3)).reshape(-1)
    crest = [dict(i=20, j=j, axis=0, z=5.4, C=1.6, ground=4.6) for j in range(ny)]
    for i in range(36, 50):
        crest += [dict(i=i, j=i - 20, axis=1, z=4.9, C=1.44, ground=4.3),
                  dict(i=i + 1, j=i - 20, axis=0, z=4.9, C=1.44, ground=4.3)]
    opts = dict(nx=nx, ny=ny, dx=dx, z=z, manning=0.03, crest=crest,
                bc={"west": "wall", "east": "outflow", "south": "wall", "north": "wall"})
    lift = [f["z"] if m < ny else f["z"] + 0.5 for m, f in enumerate(crest)]
    return opts, lift, (lambda X, Y, Z: np.where(X < 30, 6 - Z, 0))


MMHR_ = 1 / 1000 / 3600


def gi_case():
    """tests/gpu-suite.js:188-193: GI stores (100 mm/h intake, 20 mm capacity, 5 mm/h drain) on the west
    half of a 60 x 40 slope; 60 mm/h rain for 1800 s, then none until 3600 s."""
    nx, ny, dx = 60, 40, 2
    I, _ = np.meshgrid(np.arange(nx), np.arange(ny))
    z = (0.01 * I).reshape(-1)
    half = (I < nx / 2).reshape(-1)
    gi = dict(rate=np.where(half, 100 * MMHR_, 0.0), cap=np.where(half, 0.02, 0.0), drain=np.where(half, 5 * MMHR_, 0.0))
    return dict(nx=nx, ny=ny, dx=dx, z=z, manning=0.03, bc={"west": "outflow"}, gi=gi), [(60 * MMHR_, 1800.0),
                                                                                         (0.0, 3600.0)]


def ga_case():
    """tests/gpu-suite.js:332-345: Green-Ampt on a 60 x 30 slope with a valley, two soils (K 3e-6 / 1e-6 m/s,
    psi dtheta 0.03 / 0.08 m) and a shallow water table on half (Fmax 0.02 m); 120 mm/h for 1800 s, then none
    until 2700 s."""
    nx, ny, dx = 60, 30, 2
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    z = (5 - 0.03 * I + 0.06 * np.abs(J - 14.5)).reshape(-1)
    K = np.where(I < 30, 3e-6, 1e-6).reshape(-1)
    ga = dict(psiD=np.where(I < 30, 0.03, 0.08).reshape(-1), fmax=np.where(J < 15, 0.02, 1e3).reshape(-1),
              perv=np.full(nx * ny, 0.8))
    return dict(nx=nx, ny=ny, dx=dx, z=z, manning=0.04, infil=K, ga=ga,
                bc={"west": "wall", "east": "outflow", "south": "wall", "north": "wall"}), [(120 * MMHR_, 1800.0),
                                                                                              (0.0, 2700.0)]
