"""Seeded synthetic flood scenarios (terrain, roughness, rainfall, initial state).

The reference can only describe constant or planar fields through its config
(Input/FieldGenerator.cpp:114-135); these generators produce the arrays that
the solver receives instead, the way a raster would be injected.  Everything
is derived from numpy's PCG64 stream of the given seed, so a seed fixes the
bytes of every array (tests/golden/ holds SHA-256 digests per seed).

Scenario families (sizes are the caller's):

* ``tilted_hills``: plane rising 0.01 m/m along x (the x = 0 edge is lowest),
  Gaussian hills and spectral fractional-Brownian-motion noise.
* ``voronoi_classes``: Manning n from nearest-site land-use classes.
* ``storm``: a Gaussian storm cell drifting at constant velocity.
* ``dam_break``: flat bed + fBm, a wall column with a breach, 5 m upstream.
"""
from __future__ import annotations

import dataclasses
import hashlib
import math

import numpy as np

MM_PER_H = 1.0e-3 / 3600.0  # m/s per (mm/h)


def fbm(rng: np.random.Generator, ny: int, nx: int, hurst: float, amplitude: float):
    """Spectral fBm: random phases, power ~ k^-(2H+2), scaled to max |z| = amplitude."""
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(2.0 * hurst + 2.0) / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    spec = amp * np.exp(1j * phase)
    z = np.fft.irfft2(spec, s=(ny, nx))
    m = np.max(np.abs(z))
    if m > 0:
        z *= amplitude / m
    return z


def tilted_hills(seed: int, nx: int, ny: int, dx: float, n_hills: int = 8,
                 fbm_amp: float = 0.5, base: float = 0.0, slope: float = 0.01,
                 hurst: float = 0.7):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5)[None, :]
    y = (np.arange(ny) + 0.5)[:, None]
    z = base + slope * dx * x + np.zeros((ny, 1))
    for _ in range(n_hills):
        hgt = rng.uniform(2.0, 10.0)
        sig = rng.uniform(10.0, 40.0)
        cx = rng.uniform(0, nx)
        cy = rng.uniform(0, ny)
        z = z + hgt * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sig * sig))
    z = z + fbm(rng, ny, nx, hurst, fbm_amp)
    return np.ascontiguousarray(z, dtype=np.float64)


def voronoi_classes(seed: int, nx: int, ny: int, values=(0.02, 0.04, 0.06, 0.08),
                    sites_per_class: int = 6):
    rng = np.random.default_rng(seed)
    ns = len(values) * sites_per_class
    sx = rng.uniform(0, nx, ns)
    sy = rng.uniform(0, ny, ns)
    cls = np.repeat(np.arange(len(values)), sites_per_class)
    rng.shuffle(cls)
    x = np.arange(nx) + 0.5
    y = np.arange(ny) + 0.5
    best = np.full((ny, nx), np.inf)
    out = np.zeros((ny, nx))
    for k in range(ns):
        d = (y[:, None] - sy[k]) ** 2 + (x[None, :] - sx[k]) ** 2
        m = d < best
        best[m] = d[m]
        out[m] = values[cls[k]]
    return out


@dataclasses.dataclass
class Storm:
    peak: float   # m/s
    sigma: float  # m
    cx0: float
    cy0: float
    vx: float
    vy: float

    def as_dict(self):
        return dataclasses.asdict(self)

    def rate(self, x, y, t):
        cx = self.cx0 + self.vx * t
        cy = self.cy0 + self.vy * t
        return self.peak * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * self.sigma ** 2))


def drifting_storm(seed: int, nx: int, ny: int, dx: float, peak_mm_h=80.0,
                   sigma_m=400.0, speed=5.0):
    rng = np.random.default_rng(seed)
    heading = rng.uniform(0.0, 2.0 * np.pi)
    lx, ly = nx * dx, ny * dx
    cx0 = rng.uniform(0.2, 0.8) * lx
    cy0 = rng.uniform(0.2, 0.8) * ly
    return Storm(peak_mm_h * MM_PER_H, sigma_m, cx0, cy0,
                 speed * math.cos(heading), speed * math.sin(heading))


@dataclasses.dataclass
class Scenario:
    """One catchment: grid, bed, initial state, roughness, forcing, horizon."""
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    zb: np.ndarray
    h0: np.ndarray
    u0: np.ndarray
    v0: np.ndarray
    n_manning: np.ndarray | float | None
    d_shallow: float = 0.1
    d_deep: float = 0.3
    rain_t: np.ndarray | None = None      # uniform rainfall series (s, m/s)
    rain_v: np.ndarray | None = None
    storm: Storm | None = None
    head_cols: int = 0                   # parity-mode outflow: head(0) on x < head_cols
    flags: int = 0
    cn_target: float = 1.0
    h_dry: float = 1e-3                  # E8: no velocity where h <= h_dry
    u_max: float = 10.0                  # E9: velocity components within +-u_max
    # further source terms and boundaries (None = off):
    energy_loss: tuple | None = None      # (fx, fy) per metre, scalars or per-cell arrays
    infiltration: tuple | None = None     # (rate m/s, initial capacity m)
    eddy: dict | None = None              # {"mu": value/array} or {"smagorinsky": (c, s)}
    stage_t: np.ndarray | None = None     # stage boundary: water-level series (s, m)
    stage_v: np.ndarray | None = None
    stage_sel: np.ndarray | None = None   # cells of the stage boundary (ny x nx, bool)
    manning_fields: tuple | None = None   # per-cell (n_shallow, n_deep, d_shallow, d_deep)
    t_end: float = 3600.0
    step_duration: float = 60.0
    initial_dt: float = 1.0
    scheme: str = "heun2"

    @property
    def ncells(self):
        return self.nx * self.ny

    def digest(self):
        h = hashlib.sha256()
        for a in (self.zb, self.h0, self.u0, self.v0):
            h.update(np.ascontiguousarray(a, np.float64).tobytes())
        if isinstance(self.n_manning, np.ndarray):
            h.update(np.ascontiguousarray(self.n_manning, np.float64).tobytes())
        return h.hexdigest()


# Extension flag values shared with the HIP library and the oracle.
E1_RAIN_GUARD = 0x01
E2_WEST_OUTFLOW = 0x02
E4_STORM = 0x08
E7_Y_ORIENT = 0x40
E8_DRY_VELOCITY = 0x80
E9_VELOCITY_CAP = 0x100
# physics mode: rain on dry cells, free outflow, y orientation fixed, dry
# cells carry no velocity, velocity components capped
PHYSICS = 0x01 | 0x02 | 0x40 | 0x80 | 0x100


def rain_series(rate_m_s: float, duration_s: float, t_end: float, step: float):
    """Uniform rainfall on for duration_s then off, with a one-step ramp and
    samples past t_end + step (the reference reads past the last sample
    otherwise, Input/TimeSeries.cpp:32-35)."""
    t = np.array([0.0, duration_s, duration_s + step, max(t_end, duration_s + step) + 2 * step])
    v = np.array([rate_m_s, rate_m_s, 0.0, 0.0])
    return t, v


def hillslope_rain(seed=0, n=256, dx=5.0, prewet=0.0, physics=False, t_end=3600.0):
    """256x256 tilted plane + 8 hills + fBm 0.5 m, rain 50 mm/h for 30 min,
    n = 0.03, Heun-2, 1 h.  Parity mode (physics=False): the reference's
    semantics with a head(0) sink on column x = 0 as the outflow.  Physics mode:
    E1 (rain on dry cells), E2 (free outflow at x = 0), E3 (CFL 0.5)."""
    zb = tilted_hills(seed, n, n, dx, n_hills=8, fbm_amp=0.5)
    h0 = np.full((n, n), float(prewet))
    z = np.zeros((n, n))
    rt, rv = rain_series(50.0 * MM_PER_H, 1800.0, t_end, 60.0)
    sc = Scenario("hillslope_rain", n, n, dx, dx, zb, h0, z.copy(), z.copy(), 0.03,
                  rain_t=rt, rain_v=rv, t_end=t_end)
    if physics:
        sc.flags = PHYSICS
        sc.cn_target = 0.5
    else:
        sc.head_cols = 1
    return sc


def storm_catchment(seed=1, n=2048, dx=2.0, t_end=7200.0, prewet=0.0):
    """2048^2, dx 2 m, 30 hills + fBm 1 m on +300 m base; drifting storm
    (80 mm/h peak, sigma 400 m, 5 m/s); Voronoi Manning n in [0.02, 0.08];
    2 h.  Physics mode flags E1 | E2 | E4, CFL 0.5."""
    zb = tilted_hills(seed, n, n, dx, n_hills=30, fbm_amp=1.0, base=300.0)
    nm = voronoi_classes(seed + 1000, n, n)
    st = drifting_storm(seed + 2000, n, n, dx)
    z = np.zeros((n, n))
    sc = Scenario("storm_catchment", n, n, dx, dx, zb, np.full((n, n), float(prewet)), z.copy(),
                  z.copy(), nm, storm=st, t_end=t_end,
                  flags=PHYSICS | E4_STORM, cn_target=0.5)
    return sc


def breach_front(seed=2, n=4096, dx=1.0, t_end=600.0, wall_x=1000, breach=50):
    """4096^2, dx 1 m, flat bed + fBm 0.2 m, a 1-cell wall column at x = wall_x
    (bed +10 m) with a breach of `breach` cells, 5 m upstream, dry downstream,
    n = 0.025, no rain, 600 s."""
    rng = np.random.default_rng(seed)
    zb = fbm(rng, n, n, 0.7, 0.2)
    wall_x = min(wall_x, n - 2)
    zb[:, wall_x] += 10.0
    b0 = n // 2 - breach // 2
    zb[b0:b0 + breach, wall_x] -= 10.0
    h0 = np.zeros((n, n))
    h0[:, :wall_x] = 5.0
    z = np.zeros((n, n))
    return Scenario("breach_front", n, n, dx, dx, np.ascontiguousarray(zb), h0, z.copy(), z.copy(),
                    0.025, t_end=t_end, cn_target=0.5, flags=E7_Y_ORIENT | E8_DRY_VELOCITY)


def ensemble_member(seed: int, side_lo=128, side_hi=1024, t_end=3600.0):
    """One catchment of the batched ensemble: sides log-uniform per axis in
    [side_lo, side_hi], dx U[2, 10], hillslope_rain terrain, rain U[10, 100] mm/h for
    U[15, 60] min, n U[0.02, 0.06]."""
    rng = np.random.default_rng(seed)
    nx = int(round(math.exp(rng.uniform(math.log(side_lo), math.log(side_hi)))))
    ny = int(round(math.exp(rng.uniform(math.log(side_lo), math.log(side_hi)))))
    dx = float(rng.uniform(2.0, 10.0))
    rate = float(rng.uniform(10.0, 100.0)) * MM_PER_H
    dur = float(rng.uniform(15.0, 60.0)) * 60.0
    n_man = float(rng.uniform(0.02, 0.06))
    zb = tilted_hills(seed + 77, nx, ny, dx, n_hills=8, fbm_amp=0.5)
    z = np.zeros((ny, nx))
    rt, rv = rain_series(rate, dur, t_end, 60.0)
    return Scenario(f"member{seed}", nx, ny, dx, dx, zb, z.copy(), z.copy(), z.copy(), n_man,
                    rain_t=rt, rain_v=rv, t_end=t_end,
                    flags=PHYSICS, cn_target=0.5)


def dam_break_bump(nx=64, ny=48, dx=5.0, dy=5.0, seed=3, h_up=2.0):
    """Small parity case: tilted bed with a bump, water on the left third,
    a wet/dry front, walls all round."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5)[None, :] * dx
    y = (np.arange(ny) + 0.5)[:, None] * dy
    zb = 0.002 * x + 1.5 * np.exp(-((x - 0.6 * nx * dx) ** 2 + (y - 0.5 * ny * dy) ** 2)
                                  / (2 * (0.12 * nx * dx) ** 2))
    zb = zb + rng.uniform(-0.01, 0.01, size=(ny, nx))
    h0 = np.where(x < nx * dx / 3.0, h_up, 0.0) + np.zeros((ny, 1))
    u0 = np.zeros((ny, nx))
    v0 = np.zeros((ny, nx))
    return Scenario("dambreak", nx, ny, dx, dy, np.ascontiguousarray(zb), np.ascontiguousarray(h0),
                    u0, v0, 0.03, t_end=120.0, step_duration=30.0)
