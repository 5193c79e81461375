"""lbm_step_live: calls that end after a step's sphere update, deferring that
step's band work, so the next call needs no full-grid prologue.  Live calls
followed by an lbm_step that ends the run give the state of one lbm_step call
of the same total length, bit for bit; between live calls phi, the tags, the
fresh/dead lists, the packed frame and the sphere are the reference-state
run's after the same steps; entry points that need the populations, rho or u
refuse the live state."""
import numpy as np
import pytest

from gpu_util import announce
from lbm_py import lbm
from test_gpu_batch import FIELDS, SCENES, make

pytestmark = pytest.mark.gpu


def full_state(sim, b):
    return [sim.get(k).copy() for k in FIELDS], b.as_list()


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_live_calls_then_step_equal_one_call(storage):
    announce("live calls then lbm_step vs one call", storage)
    calls = [7, 1, 12]
    total = sum(calls) + 10
    sim1, b1 = make(SCENES[2])
    sim1.set_storage(storage)
    sim1.coupled(b1, total)
    ref = full_state(sim1, b1)
    sim1.close()
    sim2, b2 = make(SCENES[2])
    sim2.set_storage(storage)
    for n in calls:
        sim2.coupled_live(b2, n)
    sim2.coupled(b2, 10)
    got = full_state(sim2, b2)
    sim2.close()
    for k, x, y in zip(FIELDS, got[0], ref[0]):
        if k in ("f", "h"):
            continue  # populations of non-fluid cells are unspecified
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), k
    assert got[1] == ref[1]


def test_live_frames_equal_reference_state_frames():
    announce("live frames vs reference-state frames")
    sim1, b1 = make(SCENES[0])
    sim2, b2 = make(SCENES[0])
    for n in (5, 2, 9, 1, 3):
        sim1.coupled(b1, n)
        sim2.coupled_live(b2, n)
        assert b1.as_list() == b2.as_list(), n
        for k in ("phi", "phiPrev", "tag", "tagPrev", "usx", "gx", "lap"):
            assert np.array_equal(sim1.get(k).view(np.uint8), sim2.get(k).view(np.uint8)), (n, k)
        assert np.array_equal(sim1.fresh_cells(), sim2.fresh_cells())
        assert np.array_equal(sim1.dead_cells(), sim2.dead_cells())
        assert np.array_equal(sim1.pack(), sim2.pack())
    sim1.close()
    sim2.close()


def test_live_state_refuses_reference_state_calls():
    announce("live state guards")
    sim, b = make(SCENES[0])
    sim.coupled_live(b, 4)
    for what in (lambda: sim.get("rho"), lambda: sim.get("f"), lambda: sim.set("phi", sim.get("phi")),
                 lambda: sim.mass(), lambda: sim.check_fields(), lambda: sim.step(b),
                 lambda: sim.set_storage("f16")):
        with pytest.raises(lbm.LbmError, match="live"):
            what()
    sim.coupled(b, 1)
    assert np.isfinite(sim.get("rho")).all()
    sim.close()
