"""Batched stepping: every member of a batch ends exactly where lbm_step on it
alone ends (same kernels, per-scene deterministic reductions), whatever the
other members are, including a member whose sphere starts outside its clamp
box (which splits the batch's call into a one-step segment)."""
import numpy as np
import pytest

from gpu_util import announce, interior_fluid, pop_fluid
from lbm_py import lbm

pytestmark = pytest.mark.gpu

SCENES = [
    # (nx, ny, nz, waterline, sphere (cx, cy, cz, r, density, vy) or None)
    (40, 28, 40, 14.0, (20.0, 9.0, 20.0, 5.0, 2.0, 0.02)),
    (33, 47, 29, 30.0, (15.2, 20.3, 14.1, 3.7, 3.1, 0.0)),
    (64, 32, 48, 18.0, (30.5, 9.5, 22.0, 7.2, 1.6, 0.03)),
    (24, 24, 24, 12.0, None),
    (37, 41, 53, 20.0, (18.0, 38.0, 26.0, 4.0, 2.5, 0.0)),  # starts below its clamp box
]
FIELDS = ["phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev", "f", "h"]


def make(spec):
    nx, ny, nz, yw, sph = spec
    sim = lbm.Sim(nx, ny, nz)
    sim.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
    sim.init_flat(yw)
    b = None
    if sph is not None:
        cx, cy, cz, r, d, vy = sph
        b = lbm.make_sphere(nx, ny, nz, cx, cy, cz, r, d)
        b.vy = vy
    return sim, b


def snapshot(sim, b):
    return ([sim.get(k).copy() for k in FIELDS], None if b is None else b.as_list(), (sim.nx, sim.ny, sim.nz))


def assert_same(a, c, what):
    """Bitwise equality of every field; populations only at interior fluid cells
    (elsewhere they are unspecified and depend on where calls were split)."""
    tag = a[0][FIELDS.index("tag")]
    fluid = interior_fluid(tag, *a[2])
    for k, x, y in zip(FIELDS, a[0], c[0]):
        if k in ("f", "h"):
            q = 27 if k == "f" else 7
            x, y = pop_fluid(x, q, fluid), pop_fluid(y, q, fluid)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (what, k)
    assert a[1] == c[1], what


@pytest.mark.parametrize("split", [-1, 0, 1])
@pytest.mark.parametrize("subset", [[0, 1, 2, 3], [0, 1, 2, 3, 4]])
def test_batch_members_equal_alone(subset, split):
    """Whatever the batch's bulk/band-row split of its stream launches (one
    launch, split, or chosen by storage kind), each member ends where it ends
    alone (with the single simulation's default split)."""
    steps = 24
    announce("batch vs alone", subset, "split", split)
    alone = []
    for j in subset:
        sim, b = make(SCENES[j])
        sim.coupled(b, steps)
        alone.append(snapshot(sim, b))
        sim.close()
    batch = lbm.Batch()
    batch.set_stream_split(split)
    sims = []
    for j in subset:
        sim, b = make(SCENES[j])
        batch.add(sim, b)
        sims.append((sim, b))
    # two calls, so the second starts from the first's end state
    batch.step(steps // 2)
    batch.step(steps - steps // 2)
    for (sim, b), ref, j in zip(sims, alone, subset):
        got = snapshot(sim, b)
        # With member 4 the batch splits its first call after one step; the
        # others still match, as a segment boundary restates the same arithmetic.
        assert_same(got, ref, f"scene {j}")
    batch.close()


def test_batch_split_call_equals_single_call():
    announce("one call vs split calls")
    sim1, b1 = make(SCENES[2])
    sim1.coupled(b1, 30)
    sim2, b2 = make(SCENES[2])
    for n in (1, 7, 2, 20):
        sim2.coupled(b2, n)
    assert_same(snapshot(sim2, b2), snapshot(sim1, b1), "split")


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("calls", [(18, 19), (17, 20), (2, 6, 2, 6, 2, 6, 2, 6, 5)])
def test_split_calls_16bit(storage, calls):
    """With 16-bit storage a call split between the stream and the collide
    stores and reloads the same 16-bit values the fused step keeps, so any
    split gives the one-call state (config 0's sphere crosses cells near steps
    17-19, and the gas holds subnormal deviations)."""
    announce("16-bit split calls", storage, calls)
    total = sum(calls)
    sim1, b1 = make(SCENES[0])
    sim1.set_storage(storage)
    sim1.coupled(b1, total)
    sim2, b2 = make(SCENES[0])
    sim2.set_storage(storage)
    for n in calls:
        sim2.coupled(b2, n)
    assert_same(snapshot(sim2, b2), snapshot(sim1, b1), f"{storage} split {calls}")
    sim1.close()
    sim2.close()
