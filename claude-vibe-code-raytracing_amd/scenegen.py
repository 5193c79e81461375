"""Seeded scene generators for the benchmark configurations.

The reference has no random number generator; these scenes use SplitMix64 so
every run of a seed builds the same initial state on any machine.

config1(nx, ny, nz, seed): single large scene -- liquid in the bottom 40% with an
interface perturbed by 8 sinusoidal modes (total amplitude <= 2 cells), a sphere
of radius 24 and density 3 at 80% height (y is down, so cy = 0.2 ny) with a
seeded initial velocity |v| <= 0.02, and the parameters tau 0.6, tauPhi 0.58
(phase mobility cs7^2 (tauPhi - 1/2) = 0.02), sigma 0.01, g 1e-5, rhoL 1,
rhoG 0.1, W 4."""
import math

import numpy as np

MASK64 = (1 << 64) - 1

PARAMS = dict(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-5, rhoL=1.0, rhoG=0.1, W=4.0)


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


def waterline_modes(rng, nx, nz, base, nmodes=8, amax=0.25):
    """y_w(x, z) = base + sum_m a_m sin(2 pi (kx x/nx + kz z/nz) + theta_m)."""
    x = np.arange(nx, dtype=np.float64)[None, :]
    z = np.arange(nz, dtype=np.float64)[:, None]
    yw = np.full((nz, nx), float(base))
    modes = []
    for _ in range(nmodes):
        a = rng.uniform(0.0, amax)
        kx = rng.randint(1, 4)
        kz = rng.randint(1, 4)
        th = rng.uniform(0.0, 2 * math.pi)
        modes.append((a, kx, kz, th))
        yw = yw + a * np.sin(2 * math.pi * (kx * x / nx + kz * z / nz) + th)
    return yw, modes


def random_velocity(rng, vmax):
    u = rng.uniform(-1.0, 1.0)
    phi = rng.uniform(0.0, 2 * math.pi)
    s = math.sqrt(max(0.0, 1 - u * u))
    mag = rng.uniform(0.0, vmax)
    return mag * s * math.cos(phi), mag * u, mag * s * math.sin(phi)


def config1(nx, ny, nz, seed=1):
    """Returns (params, yw[nz, nx], sphere dict)."""
    rng = SplitMix64(seed)
    yw, modes = waterline_modes(rng, nx, nz, 0.6 * ny)
    vx, vy, vz = random_velocity(rng, 0.02)
    sphere = dict(cx=nx * 0.5, cy=0.2 * ny, cz=nz * 0.5, r=24.0, density=3.0, vx=vx, vy=vy, vz=vz)
    return dict(PARAMS), yw, sphere


def config2(count=64, seed=2):
    """Batched scenes: axes U{32..96}, liquid fill U[0.3, 0.6] with a perturbed
    interface, one sphere each (r U[3, 10], density U[1.5, 4]) at rest at a
    random position above the interface inside its clamp box; parameters as in
    config1.  Returns a list of (dims, params, yw[nz, nx], sphere dict)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx, ny, nz = rng.randint(32, 96), rng.randint(32, 96), rng.randint(32, 96)
        fill = rng.uniform(0.3, 0.6)
        yw, _ = waterline_modes(rng, nx, nz, (1.0 - fill) * ny)
        r = rng.uniform(3.0, 10.0)
        density = rng.uniform(1.5, 4.0)
        m = r + 1.5
        top = float(yw.min()) - r - 1.0  # sphere fully above the interface (y is down)
        lo_y, hi_y = m, min(top, ny - 1 - m)
        cy = rng.uniform(lo_y, hi_y) if hi_y > lo_y else lo_y
        cx = rng.uniform(m, nx - 1 - m)
        cz = rng.uniform(m, nz - 1 - m)
        out.append(((nx, ny, nz), dict(PARAMS), yw,
                    dict(cx=cx, cy=cy, cz=cz, r=r, density=density, vx=0.0, vy=0.0, vz=0.0)))
    return out


def config4(duration, rate=20.0, seed=4):
    """Served requests: Poisson arrivals at `rate` per second over `duration`
    seconds; each axis log-uniform in 32..192 cells; step budget uniform in
    500..5000; fill, sphere and parameters drawn as in config2.  Returns a list
    of (t_arrival, dims, params, yw[nz, nx], sphere dict, steps)."""
    rng = SplitMix64(seed)
    out = []
    t = 0.0
    lo, hi = math.log(32.0), math.log(192.0)
    while True:
        t += -math.log(max(1e-300, 1.0 - rng.uniform(0.0, 1.0))) / rate
        if t >= duration:
            return out
        nx, ny, nz = (int(round(math.exp(rng.uniform(lo, hi)))) for _ in range(3))
        steps = rng.randint(500, 5000)
        fill = rng.uniform(0.3, 0.6)
        yw, _ = waterline_modes(rng, nx, nz, (1.0 - fill) * ny)
        r = min(rng.uniform(3.0, 10.0), 0.2 * min(nx, ny, nz))
        density = rng.uniform(1.5, 4.0)
        m = r + 1.5
        top = float(yw.min()) - r - 1.0
        lo_y, hi_y = m, min(top, ny - 1 - m)
        cy = rng.uniform(lo_y, hi_y) if hi_y > lo_y else lo_y
        cx = rng.uniform(m, nx - 1 - m)
        cz = rng.uniform(m, nz - 1 - m)
        out.append((t, (nx, ny, nz), dict(PARAMS), yw,
                    dict(cx=cx, cy=cy, cz=cz, r=r, density=density), steps))

