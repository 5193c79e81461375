"""lbm_step_noalloc (include/lbm.h) against lbm_step: the same kernels with the
link-momentum partials in the library's per-GPU device array, so every field,
the lists and the sphere must be equal bit for bit -- on two calls in a row
(a scene, then a smaller one, which finds the larger one's partials in the
shared array), for float32 and FP16 storage, and sizes beyond the array are
refused before any work."""
import numpy as np
import pytest

from gpu_util import announce, interior_fluid, pop_fluid
from lbm_py import lbm

pytestmark = pytest.mark.gpu

FIELDS = ["phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev"]


def scene(nx, ny, nz, storage):
    sim = lbm.Sim(nx, ny, nz)
    sim.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
    sim.init_flat(ny * 0.55)
    sim.set_storage(storage)
    b = lbm.make_sphere(nx, ny, nz, nx * 0.47, ny * 0.3, nz * 0.52, min(nx, ny, nz) * 0.15, 2.5)
    b.vx, b.vy, b.vz = 0.03, 0.05, -0.02
    return sim, b


def state(sim, b):
    out = {k: sim.get(k) for k in FIELDS}
    fluid = interior_fluid(out["tag"], sim.nx, sim.ny, sim.nz)
    out["f"] = pop_fluid(sim.get("f"), 27, fluid)
    out["h"] = pop_fluid(sim.get("h"), 7, fluid)
    out["fresh"] = np.asarray(sim.fresh_cells())
    out["dead"] = np.asarray(sim.dead_cells())
    return out, b.as_list()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_equals_step_two_calls_second_smaller(storage):
    for dims, steps in (((72, 56, 64), 30), ((40, 28, 40), 25)):
        announce("lbm_step_noalloc vs lbm_step", storage, dims, steps, "steps")
        s1, b1 = scene(*dims, storage)
        s2, b2 = scene(*dims, storage)
        s1.coupled(b1, steps)
        s2.coupled_noalloc(b2, steps)
        a, sa = state(s1, b1)
        c, sc = state(s2, b2)
        assert sa == sc, dims
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), c[k].view(np.uint8)), (dims, k)
        s1.close()
        s2.close()


def test_noalloc_refuses_grid_beyond_scratch():
    announce("lbm_step_noalloc refuses 448^3")
    sim = lbm.Sim(448, 448, 448)
    with pytest.raises(lbm.LbmError, match="device scratch"):
        sim.coupled_noalloc(None, 1)
    sim.close()
