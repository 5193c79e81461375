"""Row march (lbm_set_stream_split(sim, 2)): one launch computes phi(t+1) and
streams and collides every cell outside the band box, marching each row along
z, beside a launch over the band box's cells.  Same arithmetic in the same
order as the split form (phase pass, bulk tiles, band-row tiles), so every
field, the fluid cells' populations and the sphere are equal bit for bit, for
grids whose rows do not fill a wave, whose rows do not split evenly over the
eight XCDs, with and without a sphere, and for any number of plane segments."""
import numpy as np
import pytest

from gpu_util import announce
from lbm_py import lbm, scenegen
from test_gpu_batch import FIELDS, assert_same, make, snapshot

pytestmark = pytest.mark.gpu

MARCH_SCENES = [
    # (nx, ny, nz, waterline, sphere (cx, cy, cz, r, density, vy) or None)
    (40, 28, 40, 14.0, (20.0, 9.0, 20.0, 5.0, 2.0, 0.02)),     # config 0's box, sinking sphere
    (70, 37, 45, 20.0, (35.3, 18.2, 22.6, 6.5, 2.5, -0.03)),  # rows of 70: two waves, one partial
    (128, 64, 33, 30.0, (64.0, 32.0, 16.0, 9.0, 1.5, 0.04)),  # the sphere's box meets the planes' segments
    (24, 24, 24, 12.0, None),                                 # no sphere: no band box
    (64, 48, 40, 20.0, (6.2, 7.1, 5.8, 4.5, 2.0, -0.02)),     # box clipped at three grid faces
]
CALLS = [5, 12, 1, 9]


def run(spec, mode, segments=0, rows=2):
    sim, b = make(spec)
    sim.set_stream_split(mode)
    sim.set_march_shape(rows, segments)
    for n in CALLS:
        sim.coupled(b, n)
    out = snapshot(sim, b)
    sim.close()
    return out


@pytest.mark.parametrize("rows", [2, 1])
@pytest.mark.parametrize("j", range(len(MARCH_SCENES)))
def test_march_equals_split(j, rows):
    spec = MARCH_SCENES[j]
    announce("row march vs split", spec, "rows per workgroup", rows)
    ref = run(spec, 1)
    assert_same(run(spec, 2, rows=rows), ref, ("march", spec, rows))


@pytest.mark.parametrize("rows", [2, 1])
@pytest.mark.parametrize("segments", [1, 3, 7])
def test_march_segments_do_not_change_results(segments, rows):
    spec = MARCH_SCENES[2]
    announce("row march with", segments, "segments", rows, "rows", spec)
    assert_same(run(spec, 2, segments, rows), run(spec, 1), ("segments", segments, rows))


def test_march_autotuned_shape():
    """The autotuner times the march shapes among its candidates; whatever it
    keeps, the state is that of plain calls of the same length."""
    spec = MARCH_SCENES[1]
    announce("autotuned march", spec)
    sim, b = make(spec)
    r = sim.autotune(b, 4)
    assert r["march_rows"] in (1, 2) and r["march_segments"] >= 0, r
    # 3 split modes: one launch, split, march x 3 shapes; graph steps 0 and 2
    assert r["candidates"] == (1 + 1 + 3) * 2, r
    total = r["candidates"] * (4 + 2)
    sim.coupled(b, 7)
    tuned = snapshot(sim, b)
    sim.close()
    sim, b = make(spec)
    sim.coupled(b, total)
    sim.coupled(b, 7)
    assert_same(tuned, snapshot(sim, b), "autotuned")
    sim.close()


def test_march_live_calls_and_graphs():
    """Live continuation and captured graphs replay the march launches."""
    spec = MARCH_SCENES[0]
    announce("row march, live calls and graphs", spec)
    states = []
    for mode, g in ((1, 0), (2, 0), (2, 8)):
        sim, b = make(spec)
        sim.set_stream_split(mode)
        sim.set_graph_steps(g)
        sim.coupled_live(b, 20)
        sim.coupled_live(b, 17)
        sim.coupled(b, 11)
        states.append(snapshot(sim, b))
        sim.close()
    assert_same(states[1], states[0], "live")
    assert_same(states[2], states[0], "graphs")


def test_march_256cube_config1():
    """The benchmark's scene: 256^3, config 1 (seeded interface, r = 24 sphere
    moving), 12 steps in calls of 5 and 7, march against split."""
    announce("row march vs split, 256^3 config 1")
    params, yw, sph = scenegen.config1(256, 256, 256, 1)
    outs = []
    for mode in (1, 2):
        sim = lbm.Sim(256, 256, 256)
        sim.set_params(**params)
        sim.init_waterline_map(yw)
        b = lbm.make_sphere(256, 256, 256, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
        b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
        sim.set_stream_split(mode)
        sim.coupled(b, 5)
        sim.coupled(b, 7)
        outs.append(snapshot(sim, b))
        sim.close()
    assert_same(outs[1], outs[0], "256^3")


def test_march_mode_argument_checked():
    sim = lbm.Sim(16, 16, 16)
    with pytest.raises(lbm.LbmError):
        sim.set_stream_split(3)
    with pytest.raises(lbm.LbmError):
        sim.set_march_shape(2, -1)
    with pytest.raises(lbm.LbmError):
        sim.set_march_shape(3, 0)
    sim.close()
