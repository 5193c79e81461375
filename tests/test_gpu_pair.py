"""The 16-bit stream kernel with two cells per lane (aligned 4-byte pairs)
against the one-cell kernel: the same state bit for bit -- every field, the
sphere, and the populations of every interior fluid cell -- for grids of the
three pull-offset parity classes (nx even; nx odd and ny even; nx and ny odd),
alone and in a batch, with a sphere exchanging momentum and with none."""
import numpy as np
import pytest

from gpu_util import announce, interior_fluid, pop_fluid
from lbm_py import lbm
from test_gpu_batch import FIELDS, SCENES, make

pytestmark = pytest.mark.gpu

# SCENES holds parity classes 0 (40x28x40, 64x32x48, 24^3) and 2 (33x47x29,
# 37x41x53); add class 1 (nx odd, ny even), with its sphere near a wall.
PAIR_SCENES = SCENES + [(35, 30, 26, 15.0, (9.2, 11.0, 12.5, 4.6, 2.2, 0.01))]


def state(sim, b):
    return [sim.get(k).copy() for k in FIELDS], None if b is None else b.as_list(), (sim.nx, sim.ny, sim.nz)


def assert_same(a, c, what):
    tag = a[0][FIELDS.index("tag")]
    fluid = interior_fluid(tag, *a[2])
    for k, x, y in zip(FIELDS, a[0], c[0]):
        if k in ("f", "h"):
            q = 27 if k == "f" else 7
            x, y = pop_fluid(x, q, fluid), pop_fluid(y, q, fluid)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (what, k)
    assert a[1] == c[1], what


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("j", range(len(PAIR_SCENES)))
def test_pair_kernel_equals_one_cell_kernel(storage, j):
    announce("pair vs one-cell stream kernel", storage, PAIR_SCENES[j][:3])
    out = []
    for lanes in (1, 2):
        sim, b = make(PAIR_SCENES[j])
        sim.set_storage(storage)
        sim.set_stream_lanes(lanes)
        sim.coupled(b, 17)
        sim.coupled(b, 13)
        out.append(state(sim, b))
        sim.close()
    assert_same(out[1], out[0], f"{storage} scene {j}")


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_pair_kernel_batch_equals_one_cell_alone(storage):
    announce("pair batch vs one-cell alone", storage)
    alone = []
    for spec in PAIR_SCENES:
        sim, b = make(spec)
        sim.set_storage(storage)
        sim.set_stream_lanes(1)
        sim.coupled(b, 12)
        sim.coupled(b, 12)
        alone.append(state(sim, b))
        sim.close()
    batch = lbm.Batch()
    batch.set_stream_lanes(2)
    sims = []
    for spec in PAIR_SCENES:
        sim, b = make(spec)
        sim.set_storage(storage)
        batch.add(sim, b)
        sims.append((sim, b))
    batch.step(12)
    batch.step(12)
    for j, ((sim, b), ref) in enumerate(zip(sims, alone)):
        assert_same(state(sim, b), ref, f"{storage} scene {j}")
    batch.close()
    for sim, _ in sims:
        sim.close()


def test_stream_lanes_argument_checked():
    sim = lbm.Sim(16, 16, 16)
    for bad in (0, 3, -1):
        with pytest.raises(lbm.LbmError):
            sim.set_stream_lanes(bad)
    sim.set_stream_lanes(1)
    sim.set_stream_lanes(2)
    sim.close()
