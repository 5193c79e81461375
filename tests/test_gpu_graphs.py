"""Launch graphs: the steady steps of a fused call replayed from a captured HIP
graph give the state that launching every kernel on its own gives, bit for bit
at every cell (same kernels, same arguments, same order), for one scene and for
a batch whose members change between calls."""
import numpy as np
import pytest

from gpu_util import announce
from lbm_py import lbm
from test_gpu_batch import FIELDS, SCENES, make

pytestmark = pytest.mark.gpu

# Calls of 1, 2 and 3 steps have no steady step; 11 and 17 replay graphs.
CALLS = [11, 1, 17, 2, 3, 10]


def full_state(sim, b):
    return [sim.get(k).copy() for k in FIELDS], None if b is None else b.as_list()


def assert_identical(a, c, what):
    for k, x, y in zip(FIELDS, a[0], c[0]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (what, k)
    assert a[1] == c[1], what


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("gsteps", [2, 8])
def test_single_scene_graph_equals_eager(storage, gsteps):
    announce("graph vs eager, one scene", storage, gsteps)
    states = []
    for g in (0, gsteps):
        sim, b = make(SCENES[0])
        sim.set_storage(storage)
        sim.set_graph_steps(g)
        for n in CALLS:
            sim.coupled(b, n)
        caps, reps = sim.graph_counts()
        if g == 0:
            assert (caps, reps) == (0, 0)
        else:
            assert caps >= 1 and reps >= 1, (caps, reps)
        states.append(full_state(sim, b))
        sim.close()
    assert_identical(states[1], states[0], f"{storage} graph {gsteps}")


def test_batch_graph_equals_eager_with_membership_changes():
    announce("graph vs eager, batch with membership changes")
    subset = [0, 1, 2, 3]
    results = []
    for g in (0, 4):
        batch = lbm.Batch()
        batch.set_graph_steps(g)
        sims = [make(SCENES[j]) for j in subset]
        for sim, b in sims:
            batch.add(sim, b)
        batch.step(12)
        batch.remove(sims[1][0])   # different launch grids: a new graph
        batch.step(9)
        batch.add(*sims[1])        # the first grids again: the cached graph
        batch.step(12)
        caps, reps = batch.graph_counts()
        if g:
            assert caps == 2 and reps >= 4, (caps, reps)
        results.append([full_state(sim, b) for sim, b in sims])
        batch.close()
        for sim, _ in sims:
            sim.close()
    for j, (x, y) in enumerate(zip(results[1], results[0])):
        assert_identical(x, y, f"scene {subset[j]}")


def test_graph_steps_argument_checked():
    sim = lbm.Sim(16, 16, 16)
    for bad in (1, 3, -2, 66):
        with pytest.raises(lbm.LbmError):
            sim.set_graph_steps(bad)
    sim.set_graph_steps(0)
    sim.set_graph_steps(64)
    sim.close()


def test_autotune_steps_like_plain_calls():
    """The autotuner's trials are ordinary steps: afterwards every member is
    where plain calls of the same total length leave it (interior fluid
    populations and every other field, bit for bit)."""
    from test_gpu_batch import assert_same, snapshot
    announce("autotune vs plain calls")
    subset = [0, 1, 2]
    tuned = lbm.Batch()
    sims = [make(SCENES[j]) for j in subset]
    for sim, b in sims:
        sim.set_storage("f16")
        tuned.add(sim, b)
    r = tuned.autotune(6)
    assert r["cells_per_lane"] in (1, 2) and r["graph_steps"] in (0, 2, 8, 32)
    # 2 lane counts x split or not x graph steps 0, 2
    assert r["candidates"] == 8 and r["ms_per_step"] > 0, r
    assert r["stream_split"] in (0, 1), r
    total = r["candidates"] * (6 + 2)
    tuned.step(5)
    plain = lbm.Batch()
    refs = [make(SCENES[j]) for j in subset]
    for sim, b in refs:
        sim.set_storage("f16")
        plain.add(sim, b)
    plain.step(total)
    plain.step(5)
    for j, ((s1, b1), (s2, b2)) in enumerate(zip(sims, refs)):
        assert_same(snapshot(s1, b1), snapshot(s2, b2), f"scene {subset[j]}")
    for bt in (tuned, plain):
        bt.close()
    for sim, _ in sims + refs:
        sim.close()


def test_capture_beside_create_and_destroy_on_another_thread():
    """One thread steps fresh simulations whose first call captures a graph
    while another thread creates and destroys simulations (device allocations,
    releases, streams): every call succeeds and each result equals eager
    launches bit for bit."""
    import threading
    announce("graph capture beside create/destroy on another thread")
    sim, b = make(SCENES[0])
    sim.set_graph_steps(0)
    sim.coupled(b, 12)
    ref = full_state(sim, b)
    sim.close()
    stop = threading.Event()
    churn = [0]
    errors = []

    def churner():
        try:
            while not stop.is_set():
                s = lbm.Sim(48, 40, 44)
                s.close()
                churn[0] += 1
        except Exception as e:  # noqa: BLE001 (reported below)
            errors.append(e)

    th = threading.Thread(target=churner)
    th.start()
    try:
        for rep in range(12):
            sim, b = make(SCENES[0])
            sim.set_graph_steps(2)
            sim.coupled(b, 12)
            caps, reps = sim.graph_counts()
            assert caps == 1 and reps >= 1, (rep, caps, reps)
            assert_identical(full_state(sim, b), ref, f"capture beside churn {rep}")
            sim.close()
    finally:
        stop.set()
        th.join()
    assert not errors, errors
    assert churn[0] > 0
    announce("simulations created and destroyed meanwhile:", churn[0])
