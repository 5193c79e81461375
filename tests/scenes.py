"""Scene set-ups shared by the tests: each mirrors one reference test script or
one benchmark configuration.  `apply(sim, sphere_factory)` works for both the
oracle (tests/oracle_py.py) and the HIP library binding, which expose the same
method names."""

W7 = [1 / 4, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8]


def smoke12(sim, mk):
    """test/smoke.mjs:7-19: initFlat(6) before the parameters are set."""
    sim.init_flat(6)
    sim.set_params(tau=0.7, tauPhi=0.7, sigma=0.012, gravity=0.0012, rhoL=1.0, rhoG=0.1, W=4.0)
    return mk(12, 12, 12, 6, 3, 6, 2.0, 1.5)


def config0(sim, mk):
    """40x28x40: parameters first, initFlat(14), sphere r=5 density 2 at (20,22,20)."""
    sim.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-5, rhoL=1.0, rhoG=0.1, W=4.0)
    sim.init_flat(14)
    return mk(40, 28, 40, 20, 22, 20, 5, 2.0)


def bench40(sim, mk):
    """test/bench.mjs:6-11: initFlat(14) with default parameters, then parameters."""
    sim.init_flat(14)
    sim.set_params(tau=0.7, tauPhi=0.7, sigma=0.012, gravity=0.00135, rhoL=1.0, rhoG=0.1, W=4.0)
    return mk(40, 28, 40, 20, 6, 20, 4, 1.6)


def drag20(sim, mk):
    """test/drag_stress.mjs pattern on 20x16x20 with the spring toward (14,3,9)."""
    sim.set_params(tau=0.8, tauPhi=0.7, sigma=0.008, gravity=0.0015, rhoL=1.0, rhoG=0.1, W=4.0)
    sim.init_flat(16 * 0.25)
    b = mk(20, 16, 20, 10, 16 * 0.65, 10, 3, 1.4)
    b.dragging = 1.0
    b.dragTargetX, b.dragTargetY, b.dragTargetZ = 14.0, 3.0, 9.0
    return b


def dam_fields(nx, ny, nz):
    """worker.js:54-79: phi column x < 0.35 nx, y > 0.35 ny; returns (phi, h) in
    the reference layout for the interior cells only, as index lists."""
    cells = []
    for z in range(1, nz - 1):
        for y in range(1, ny - 1):
            for x in range(1, nx - 1):
                i = (z * ny + y) * nx + x
                p = 1.0 if (x < nx * 0.35 and y > ny * 0.35) else 0.0
                cells.append((i, p))
    return cells
