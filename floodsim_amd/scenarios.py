"""Synthetic workloads for the shallow-water step (build-owned: the reference has no such generator).

urban(n, seed)          the synthetic urban tile: 0.5 % plane + fBm (Hurst 0.7, 2 m), ~30 % of the
                        cells in random axis-aligned buildings (sides 5-40 m) raised 10 m as walls,
                        a paved/pervious mask setting Manning n to 0.013 or 0.035.
dam_break(n)            256x256 dam break: flat bed, 1 m for x < n/2 m, 0.1 m elsewhere, n = 0.03.
triangle_hydrograph     0 -> peak over t_rise, back to 0 over t_fall (point inflow, m^3/s).
peaked_hyetograph       the served storm: 120 mm/h peak at 40 % of 30 min, 25 mm.
urban_inflow(n, seed)   the urban tile with a point inflow at the west edge (the large single domain).
ensemble(count, seed)   uneven batched instances: sides round(128 * 8^u), u ~ U(0, 1).
poisson_arrivals        request times of a Poisson stream.

Every array is float64, row-major (j*nx + i), 1 m cells unless stated; the same seed gives the
same arrays on every machine (numpy PCG64). `checksum` hashes the generated layers.
"""
import hashlib
import math

import numpy as np

MMHR = 1 / 1000 / 3600


def _value_noise(rng, ny, nx, wavelength):
    """Bilinear value noise on a lattice of the given spacing (cells), values in [-1, 1]."""
    ly = int(math.ceil(ny / wavelength)) + 2
    lx = int(math.ceil(nx / wavelength)) + 2
    lat = rng.uniform(-1.0, 1.0, size=(ly, lx)).astype(np.float32)
    x = (np.arange(nx, dtype=np.float32) + 0.5) / wavelength
    y = (np.arange(ny, dtype=np.float32) + 0.5) / wavelength
    ix = np.floor(x).astype(np.int64)
    fx = (x - ix).astype(np.float32)
    iy = np.floor(y).astype(np.int64)
    fy = (y - iy).astype(np.float32)
    # smoothstep weights keep the gradient continuous across lattice lines
    fx = fx * fx * (3 - 2 * fx)
    fy = fy * fy * (3 - 2 * fy)
    rows = lat[:, ix] * (1 - fx) + lat[:, ix + 1] * fx          # (ly, nx)
    out = rows[iy] * (1 - fy)[:, None] + rows[iy + 1] * fy[:, None]
    return out


def fbm(rng, ny, nx, hurst=0.7, amplitude=2.0, base=1024.0, finest=2.0):
    """Fractional Brownian surface by octave value noise: octave o has wavelength base/2^o (m at
    1 m cells) and weight 2^(-hurst*o); the sum is scaled so its bound is +-amplitude."""
    field = np.zeros((ny, nx), dtype=np.float32)
    total = 0.0
    o = 0
    lam = base
    while lam >= finest:
        w = 2.0 ** (-hurst * o)
        field += np.float32(w) * _value_noise(rng, ny, nx, lam)
        total += w
        o += 1
        lam = base / 2 ** o
    field *= np.float32(amplitude / total)
    return field


def urban(n, seed=1, ny=None, building_share=0.30, slope=0.005, amplitude=2.0, hurst=0.7, dx=1.0):
    """Synthetic urban tile (n x ny at dx m). Returns dict(nx, ny, dx, z, mask, manning)."""
    nx = int(n)
    ny = int(ny or n)
    rng = np.random.default_rng(seed)
    x = (np.arange(nx, dtype=np.float32) + 0.5) * dx
    z = fbm(rng, ny, nx, hurst=hurst, amplitude=amplitude)
    z += (slope * (nx * dx - x))[None, :].astype(np.float32)  # falls toward +x
    mask = np.ones((ny, nx), dtype=np.uint8)
    target = building_share * nx * ny
    covered = 0
    while covered < target:
        batch = 256
        w = rng.integers(5, 41, size=batch)
        h = rng.integers(5, 41, size=batch)
        x0 = rng.integers(0, max(1, nx - 4), size=batch)
        y0 = rng.integers(0, max(1, ny - 4), size=batch)
        for k in range(batch):
            sl = mask[y0[k]:y0[k] + h[k], x0[k]:x0[k] + w[k]]
            covered += int(sl.sum())
            sl[...] = 0
            if covered >= target:
                break
    z = z.astype(np.float64)
    z[mask == 0] += 10.0
    # paved (0.013) / pervious (0.035) patches: smooth noise at ~60 m, about half paved
    pav = _value_noise(rng, ny, nx, 60.0) > 0.0
    manning = np.where(pav, 0.013, 0.035)
    return dict(nx=nx, ny=ny, dx=dx, z=z.reshape(-1), mask=mask.reshape(-1), manning=manning.reshape(-1),
                seed=seed)


def dam_break(n=256):
    """Workload 1 (the reference's CPU-runnable case): flat 1 m grid, 1 m deep for x < n/2 m,
    0.1 m elsewhere, n = 0.03, walls."""
    return dict(nx=n, ny=n, dx=1.0, z=np.zeros(n * n), manning=0.03,
                depth=lambda X, Y, Z: np.where(X < n / 2, 1.0, 0.1))


def checksum(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def triangle_hydrograph(t, peak=40.0, t_rise=1200.0, t_fall=2400.0):
    """m^3/s at time t: 0 -> peak over t_rise, then back to 0 over t_fall."""
    t = np.asarray(t, dtype=np.float64)
    up = peak * t / t_rise
    down = peak * (1 - (t - t_rise) / t_fall)
    return np.where(t < 0, 0.0, np.where(t <= t_rise, up, np.maximum(0.0, down)))


def peaked_hyetograph(t, duration=1800.0, peak_mmh=120.0, peak_at=0.4, depth_mm=25.0):
    """Rain rate (m/s) of the served storm: a triangle reaching peak_mmh at peak_at * duration and holding
    depth_mm. A triangle holds peak x base / 2, so it rains for base = 2 depth / peak (25 mm at 120 mm/h:
    1500 s of the 1800 s storm), rising over peak_at of the base and falling over the rest, and the storm is
    dry before and after (25 mm, 120 mm/h, 30 min: rain from 120 s to 1620 s, peak at 720 s)."""
    t = np.asarray(t, dtype=np.float64)
    tp = peak_at * duration
    base = 2.0 * depth_mm / peak_mmh * 3600.0
    t0, t1 = tp - peak_at * base, tp + (1.0 - peak_at) * base
    if t0 < 0 or t1 > duration:
        raise ValueError("the storm's triangle does not fit its duration")
    r = np.where(t <= tp, peak_mmh * (t - t0) / (tp - t0), peak_mmh * (t1 - t) / (t1 - tp))
    return np.clip(r, 0, None) * MMHR


def urban_inflow(n, seed=2026):
    """The urban tile with a point inflow: the generator at n x n, the inflow cell at the middle of the
    west edge (its approach, 5 x 5 cells, kept open and no higher than the bed 5 cells in), all edges
    free outfall, rain 50 mm/h; the inflow follows triangle_hydrograph (0 -> 40 m^3/s over 20 min, back
    to 0 over 40 min)."""
    u = urban(n, seed=seed)
    mask = u["mask"].reshape(n, n)
    jin = n // 2
    mask[jin - 2:jin + 3, 0:5] = 1
    z = u["z"].reshape(n, n)
    z[jin - 2:jin + 3, 0:5] = np.minimum(z[jin - 2:jin + 3, 0:5], z[jin, 5])
    return dict(nx=n, ny=n, dx=1.0, z=u["z"], mask=u["mask"], manning=u["manning"],
                bc={k: "outflow" for k in ("west", "east", "south", "north")},
                inlets=[jin * n], rain=50 * MMHR, seed=seed,
                checksum=checksum(u["z"], u["mask"], u["manning"]))


def ensemble(count=256, seed=7, smin=128, smax=1024):
    """Uneven instances: side = round(smin * (smax/smin)^u); per instance a seed, a lognormal
    rain multiplier (sigma 0.3) and a Manning scale U(0.8, 1.2)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, size=count)
    sides = np.round(smin * (smax / smin) ** u).astype(int)
    seeds = rng.integers(1, 2 ** 31 - 1, size=count)
    rain_mult = rng.lognormal(0.0, 0.3, size=count)
    n_scale = rng.uniform(0.8, 1.2, size=count)
    return [dict(side=int(s), seed=int(sd), rain_mult=float(r), n_scale=float(m))
            for s, sd, r, m in zip(sides, seeds, rain_mult, n_scale)]


def poisson_arrivals(rate, count, seed=11, sides=(256, 512, 1024, 2048, 4096), probs=(0.35, 0.3, 0.2, 0.1, 0.05)):
    """Served stream: arrival times (s) at `rate` per second, a side per request, a depth scale."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    times = np.cumsum(gaps)
    sd = rng.choice(np.asarray(sides), size=count, p=np.asarray(probs))
    scale = rng.lognormal(0.0, 0.25, size=count)
    seeds = rng.integers(1, 2 ** 31 - 1, size=count)
    return [dict(t=float(t), side=int(s), depth_scale=float(c), seed=int(q))
            for t, s, c, q in zip(times, sd, scale, seeds)]
