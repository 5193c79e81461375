"""The benchmark configurations that are not single small scenes, against the
oracle (float32 storage, FP64 arithmetic: the reference's semantics).

* Config 2 as one batch: the 64 seeded scenes (axes 32..96, one sphere each)
  stepped together by lbm_batch_step in three calls, with the spheres at rest
  as generated and with seeded velocities, every member against the
  oracle's OpenMP build after each call -- solid mask and previous mask bit for
  bit, the fresh and dead cell lists equal to the ones the oracle's masks
  define, phi/rho/u within 1e-5 of the field maximum over interior fluid cells,
  the sphere centre within 1e-4 cells.
* Config 1 at 384^3: the fused float32 step against the oracle's OpenMP build
  over calls of 1 and 3 steps with the same checks, plus full-size properties
  (finite fields, no solid cell outside the interior, a nonempty solid mask).
* Served frames (config 4's loop) against the oracle's packVolume of the same
  scene stepped on its own: G (solid mask) bit for bit, R (phi * 255 truncated)
  within one count, the frame's sphere within 1e-4 cells.
"""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import announce, interior_fluid, rel_err
from lbm_py import lbm, scenegen

pytestmark = pytest.mark.gpu

TAG_FLUID, TAG_SOLID = 0, 1
FIELDS = ("phi", "rho", "ux", "uy", "uz")


def interior_mask(nx, ny, nz):
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.ravel()


def gpu_state(sim, b):
    """What the checks read: tags, macroscopic fields, lists, sphere centre."""
    return dict(tag=sim.get("tag"), tagPrev=sim.get("tagPrev"),
                fields={k: sim.get(k) for k in FIELDS},
                fresh=np.asarray(sim.fresh_cells(), np.int64),
                dead=np.asarray(sim.dead_cells(), np.int64),
                centre=(b.cx, b.cy, b.cz))


def check_against_oracle(got, so, bo, dims, what):
    nx, ny, nz = dims
    c = got["centre"]
    d = max(abs(bo.cx - c[0]), abs(bo.cy - c[1]), abs(bo.cz - c[2]))
    assert d <= 1e-4, (what, "sphere", d)
    assert np.array_equal(got["tag"], so.tag), (what, "tag")
    assert np.array_equal(got["tagPrev"], so.tagPrev), (what, "tagPrev")
    inner = interior_mask(nx, ny, nz)
    fresh = np.flatnonzero(inner & (so.tag == TAG_FLUID) & (so.tagPrev == TAG_SOLID))
    dead = np.flatnonzero(inner & (so.tag == TAG_SOLID) & (so.tagPrev == TAG_FLUID))
    assert np.array_equal(got["fresh"], fresh), (what, "fresh")
    assert np.array_equal(got["dead"], dead), (what, "dead")
    mask = interior_fluid(so.tag, nx, ny, nz)
    worst = 0.0
    for k in FIELDS:
        e = rel_err(got["fields"][k], getattr(so, k), mask)
        assert e <= 1e-5, (what, k, e)
        worst = max(worst, e)
    return worst, fresh.size, dead.size


# The first call's single step has the first retag's dead cells (every sphere
# cell); the later calls end where spheres moving across cells have both kinds.
BATCH_CALLS = (1, 19, 30)


def batch_velocities(scenes, moving):
    """Initial sphere velocities: config 2's spheres start at rest; the moving
    variant gives each a seeded velocity of at most 0.05 cells per step, so
    every sphere crosses cells within the run."""
    if not moving:
        return [(0.0, 0.0, 0.0)] * len(scenes)
    rng = scenegen.SplitMix64(5)
    return [scenegen.random_velocity(rng, 0.05) for _ in scenes]


@pytest.mark.parametrize("moving", [False, True])
def test_config2_batch_of_64_vs_oracle(moving):
    scenes = scenegen.config2(64, seed=2)
    vel = batch_velocities(scenes, moving)
    announce("config 2: 64 seeded scenes in one batch, calls of", BATCH_CALLS, "steps,",
             "moving spheres" if moving else "spheres at rest")
    batch = lbm.Batch()
    members = []
    for ((nx, ny, nz), params, yw, sph), v in zip(scenes, vel):
        sim = lbm.Sim(nx, ny, nz)
        sim.set_params(**params)
        sim.init_waterline_map(yw)
        b = lbm.make_sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
        b.vx, b.vy, b.vz = v
        batch.add(sim, b)
        members.append((sim, b))
    snaps = []
    for n in BATCH_CALLS:
        batch.step(n)
        snaps.append([gpu_state(sim, b) for sim, b in members])
        announce("batch call of", n, "steps done")
    batch.close()
    for sim, _ in members:
        sim.close()
    worst, flips = 0.0, [0, 0]
    for j, (dims, params, yw, sph) in enumerate(scenes):
        nx, ny, nz = dims
        so = O.OSim(nx, ny, nz, bits="32_omp")
        so.set_params(**params)
        so.init_waterline_map(yw)
        bo = O.sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"],
                      bits="32_omp")
        bo.vx, bo.vy, bo.vz = vel[j]
        for c, n in enumerate(BATCH_CALLS):
            so.coupled(bo, n)
            e, nf, nd = check_against_oracle(snaps[c][j], so, bo, dims, (j, dims, c))
            worst = max(worst, e)
            flips[0] += nf
            flips[1] += nd
        del so
    announce("64 members match the oracle; worst field error", worst, "fresh/dead cells", flips)
    assert flips[1] > 0, flips
    if moving:
        assert flips[0] > 0, flips


N384 = 384


def scene384(bits=None):
    params, yw, sph = scenegen.config1(N384, N384, N384, 1)
    if bits is None:
        sim = lbm.Sim(N384, N384, N384)
        b = lbm.make_sphere(N384, N384, N384, sph["cx"], sph["cy"], sph["cz"], sph["r"],
                            sph["density"])
    else:
        sim = O.OSim(N384, N384, N384, bits=bits)
        b = O.sphere(N384, N384, N384, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"],
                     bits=bits)
    sim.set_params(**params)
    sim.init_waterline_map(yw)
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    return sim, b


def test_fused_vs_oracle_384cube():
    announce("fused float32 vs oracle (OpenMP build), 384^3 config-1 scene, calls of 1 and 3")
    sg, bg = scene384()
    snaps = []
    for n in (1, 3):
        sg.coupled(bg, n)
        st = gpu_state(sg, bg)
        # Full-size property: finite macroscopic fields.
        assert all(np.isfinite(v).all() for v in st["fields"].values())
        snaps.append(st)
    sg.close()
    so, bo = scene384("32_omp")
    inner = interior_mask(N384, N384, N384)
    for c, n in enumerate((1, 3)):
        so.coupled(bo, n)
        announce("384^3 after a call of", n)
        check_against_oracle(snaps[c], so, bo, (N384, N384, N384), ("384", c))
        solid = snaps[c]["tag"] == TAG_SOLID
        assert not solid[~inner].any()
        assert solid.sum() > 0


def test_served_frames_vs_oracle_pack():
    """A served request's frames (src/worker.js:94-111: packVolume after each
    frame interval) against the oracle stepping the same scene on its own."""
    scenes = scenegen.config2(64, seed=2)
    order = sorted(range(len(scenes)), key=lambda i: np.prod(scenes[i][0]))
    (nx, ny, nz), params, yw, sph = scenes[order[5]]
    announce("served frames vs oracle packVolume", (nx, ny, nz))
    srv = lbm.Server(4 << 30)
    b = lbm.make_sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
    rid = srv.submit(nx, ny, nz, params, 60, frame_every=10, waterline=yw, sphere=b)
    frames = []
    while True:
        fr = srv.next_frame(timeout_ms=60000)
        assert fr is not None, srv.stats()
        info, rg8 = fr
        assert info.status == 0 and info.request == rid
        frames.append((info.step, rg8.copy(), (info.sphere.cx, info.sphere.cy, info.sphere.cz)))
        if info.last:
            break
    srv.close()
    assert [f[0] for f in frames] == [10, 20, 30, 40, 50, 60]
    so = O.OSim(nx, ny, nz)
    so.set_params(**params)
    so.init_waterline_map(yw)
    bo = O.sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
    done = 0
    for step, rg8, centre in frames:
        so.coupled(bo, step - done)
        done = step
        ref = so.pack()
        r_got, g_got = rg8[0::2].astype(np.int32), rg8[1::2]
        r_ref, g_ref = ref[0::2].astype(np.int32), ref[1::2]
        assert np.array_equal(g_got, g_ref), step
        assert np.max(np.abs(r_got - r_ref)) <= 1, step
        d = max(abs(bo.cx - centre[0]), abs(bo.cy - centre[1]), abs(bo.cz - centre[2]))
        assert d <= 1e-4, (step, d)
