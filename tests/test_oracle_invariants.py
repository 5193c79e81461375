"""Checks of the oracle where the reference's vectors leave it open, against
what the mathematics fixes."""
import math

import numpy as np

import oracle_py as O


def test_waterline_map_constant_equals_init_flat():
    a = O.OSim(10, 9, 8)
    b = O.OSim(10, 9, 8)
    for s in (a, b):
        s.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
    a.init_flat(4.3)
    b.init_waterline_map(np.full(8 * 10, 4.3))
    for k in ("f", "h", "phi", "phiPrev", "rho", "tag"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_waterline_map_closed_form():
    nx, ny, nz = 7, 11, 5
    s = O.OSim(nx, ny, nz)
    s.set_params(W=3.0, gravity=2e-4, rhoL=1.0, rhoG=0.2)
    rng = np.random.default_rng(5)
    yw = rng.uniform(3, 8, size=(nz, nx))
    s.init_waterline_map(yw.ravel())
    phi = s.phi.reshape(nz, ny, nx)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                want = np.float32(0.5 * (1 + math.tanh(2 * (y - yw[z, x]) / 3.0)))
                assert phi[z, y, x] == want
    # hydrostatic seed: rho(y) = 1 + 3 g sum_{y'<y} rho_phi(y') per column
    rho = s.rho.reshape(nz, ny, nx)
    for z in range(nz):
        for x in range(nx):
            p = 1.0
            for y in range(ny):
                assert abs(rho[z, y, x] - p) <= 1e-6
                p += (0.2 + float(phi[z, y, x]) * 0.8) * 2e-4 * 3.0


def test_equilibrium_moments_single_cell():
    """One interior cell of a uniform resting state collides to itself, and the
    Guo source has zero mass: sum_k S_k = 0 (closed form of sim.js:457)."""
    s = O.OSim(5, 5, 5)
    s.set_params(tau=0.7, tauPhi=0.7, sigma=0.0, gravity=0.0, rhoL=1.0, rhoG=1.0, W=4.0)
    s.init_flat(2.0)
    s.gradients()
    s.collide_hydro()
    i = (2 * 5 + 2) * 5 + 2
    f = s.f[i * 27:(i + 1) * 27].astype(np.float64)
    f2 = s.f2[i * 27:(i + 1) * 27].astype(np.float64)
    assert np.max(np.abs(f2 - f)) <= 1e-7
