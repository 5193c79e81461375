"""The serving loop (lbm_server_*): continuous batching of resident scenes
with RG8 frames.  A request's frames are the same alone and in a mix of other
requests (bit for bit, RG8 volume and sphere state); admission respects the
memory budget; frames equal Sim.packVolume of the same scene stepped on its
own."""
import time

import numpy as np
import pytest

from gpu_util import announce
from lbm_py import lbm, scenegen

pytestmark = pytest.mark.gpu

PARAMS = dict(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-5, rhoL=1.0, rhoG=0.1, W=4.0)


_SCENES = []


def scene(idx):
    """Small seeded scenes: (dims, yw, sphere)."""
    if not _SCENES:
        sc = scenegen.config2(64, seed=2)
        order = sorted(range(len(sc)), key=lambda i: np.prod(sc[i][0]))
        _SCENES.extend(sc[i] for i in order[:8])
    (nx, ny, nz), params, yw, s = _SCENES[idx]
    b = lbm.make_sphere(nx, ny, nz, s["cx"], s["cy"], s["cz"], s["r"], s["density"])
    return (nx, ny, nz), yw, b


def submit(srv, idx, steps, storage="f32", frame_every=10):
    (nx, ny, nz), yw, b = scene(idx)
    return srv.submit(nx, ny, nz, PARAMS, steps, frame_every=frame_every, waterline=yw, sphere=b,
                      storage=storage)


def drain(srv, ids, timeout_s=120):
    """Frames of the given requests until each has sent its last one."""
    got = {i: [] for i in ids}
    open_ids = set(ids)
    waited = 0.0
    while open_ids:
        fr = srv.next_frame(timeout_ms=1000)
        if fr is None:
            waited += 1.0
            assert waited < timeout_s, ("no frames", srv.stats())
            continue
        info, rg8 = fr
        assert info.status == 0, info.status
        if info.request in got:
            got[info.request].append((info.index, info.step, info.last, rg8,
                                      tuple(info.sphere.as_list()), info.latency))
            if info.last:
                open_ids.discard(info.request)
    return got


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_serve_same_alone_and_in_mix(storage):
    announce("serve alone vs mix", storage)
    budget = 8 << 30
    srv = lbm.Server(budget)
    a = submit(srv, 3, 60, storage)
    alone = drain(srv, [a])[a]
    srv.close()
    srv = lbm.Server(budget)
    srv.pause()
    b = submit(srv, 0, 120, storage)
    c = submit(srv, 1, 150, storage)
    srv.pause(False)
    # A and D arrive while B and C are on the GPU
    t0 = time.time()
    while srv.stats()["frames"] == 0:
        assert time.time() - t0 < 60, srv.stats()
        time.sleep(0.001)
    srv.pause()
    a2 = submit(srv, 3, 60, storage)
    d = submit(srv, 2, 30, storage, frame_every=7)
    srv.pause(False)
    mix = drain(srv, [b, c, a2, d])
    st = srv.stats()
    srv.close()
    assert st["max_active_seen"] >= 3, st
    got = mix[a2]
    assert [f[:3] for f in got] == [f[:3] for f in alone]
    assert len(alone) == 6 and alone[-1][2] == 1
    for x, y in zip(got, alone):
        assert np.array_equal(x[3], y[3]), x[1]
        assert x[4] == y[4], x[1]
    # every request ended with its budget of steps and a frame per interval
    assert [f[1] for f in mix[d]] == [7, 14, 21, 28, 30]


def test_serve_admission_by_budget():
    announce("serve admission by budget")
    sizes = [np.prod(scene(i)[0]) for i in range(3)]
    dims = [scene(i)[0] for i in range(3)]
    need = max(lbm.state_bytes(*d) for d in dims)
    srv = lbm.Server(need + need // 4)  # room for one scene at a time
    ids = [submit(srv, i, 20) for i in range(3)]
    got = drain(srv, ids)
    st = srv.stats()
    srv.close()
    assert st["max_active_seen"] == 1, st
    assert st["completed"] == 3 and st["admitted"] == 3, st
    for i in ids:
        assert [f[1] for f in got[i]] == [10, 20]
    assert st["lattice_updates"] == 20 * sum(sizes)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_serve_scene_in_reused_memory(storage):
    """A server's scenes take device memory that finished scenes gave back
    (its memory pool): a request served after others have come and gone, one
    at a time, sends the frames it sends on a fresh server."""
    announce("serve scene in reused memory", storage)
    dims = [scene(i)[0] for i in (3, 4, 5)]
    need = max(lbm.state_bytes(*d, storage) for d in dims)
    srv = lbm.Server(4 << 30)
    a = submit(srv, 3, 40, storage)
    fresh = drain(srv, [a])[a]
    srv.close()
    # One scene at a time, the larger ones first, so the last one's buffers
    # come from memory they wrote.
    srv = lbm.Server(need + need // 4)
    ids = [submit(srv, i, 20, storage) for i in (5, 4)]
    a2 = submit(srv, 3, 40, storage)
    got = drain(srv, ids + [a2])
    st = srv.stats()
    srv.close()
    assert st["max_active_seen"] == 1 and st["completed"] == 3, st
    assert [f[:3] for f in got[a2]] == [f[:3] for f in fresh]
    for x, y in zip(got[a2], fresh):
        assert np.array_equal(x[3], y[3]), x[1]
        assert x[4] == y[4], x[1]


def test_serve_rejects_oversized_request():
    announce("serve rejects oversized")
    srv = lbm.Server(1 << 20)
    with pytest.raises(lbm.LbmError):
        submit(srv, 0, 10)
    srv.close()


def test_serve_frames_equal_pack_volume():
    announce("serve frames vs packVolume")
    srv = lbm.Server(4 << 30)
    a = submit(srv, 4, 30)
    frames = drain(srv, [a])[a]
    srv.close()
    (nx, ny, nz), yw, b = scene(4)
    sim = lbm.Sim(nx, ny, nz)
    sim.set_params(**PARAMS)
    sim.init_waterline_map(yw)
    for idx, step, last, rg8, sph, lat in frames:
        sim.coupled(b, 10)
        assert np.array_equal(sim.pack(), rg8), step
        assert tuple(b.as_list()) == sph, step
    sim.close()
