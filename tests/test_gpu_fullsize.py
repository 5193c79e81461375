"""The 256^3 scene the benchmark steps (perturbed waterline at 0.6 of the
height, an r = 24 sphere of density 3 moving at 0.02 cells per step), at its
full size.

float32 populations: the fused step against the oracle's OpenMP build (the
same results as its sequential build) over calls of 1, 4, 10 and 25 steps --
after every call the solid mask and the previous mask bit for bit, the fresh
and dead cell lists equal to the ones the oracle's masks define (interior
cells, ascending index, as handleFreshCells and absorbDeadCells visit them,
sim.js:257-310 and solid.js), phi/rho/u within 1e-5 of the field maximum over
interior fluid cells and the sphere centre within 1e-4 cells.

FP16 and BF16 populations, at 256^3 and 384^3: the two-cells-per-lane stream
kernel against the one-cell kernel, every field, the populations of every
interior fluid cell and the sphere bit for bit after two calls.

Every storage form against the oracle with FP64 storage and arithmetic (its
OpenMP build) after 200 steps (LBM_FULLSIZE_PREC_STEPS sets another count):
phi within the bounds the 2000-step check of the batched scenes uses, sphere
centre within 0.05 cells over the trajectory."""
import os

import numpy as np
import pytest

import oracle_py as O
from gpu_util import announce, interior_fluid, pop_fluid, rel_err
from test_gpu_precision import PHI_MAX, errors
from lbm_py import lbm, scenegen

pytestmark = pytest.mark.gpu

N = 256
SEED = 1
TAG_FLUID, TAG_SOLID = 0, 1


def gpu_scene(storage, n=N):
    params, yw, sph = scenegen.config1(n, n, n, SEED)
    sim = lbm.Sim(n, n, n)
    sim.set_params(**params)
    sim.init_waterline_map(yw)
    sim.set_storage(storage)
    b = lbm.make_sphere(n, n, n, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    return sim, b


def oracle_scene(bits="32_omp"):
    params, yw, sph = scenegen.config1(N, N, N, SEED)
    s = O.OSim(N, N, N, bits=bits)
    s.set_params(**params)
    s.init_waterline_map(yw)
    b = O.sphere(N, N, N, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"],
                 bits=bits)
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    return s, b


def interior():
    m = np.zeros((N, N, N), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.ravel()


def test_fused_vs_oracle_256cube():
    announce("fused float32 vs oracle (OpenMP build), 256^3 config-1 scene, 40 steps")
    so, bo = oracle_scene()
    sg, bg = gpu_scene("f32")
    inner = interior()
    flips = [0, 0]
    for steps in (1, 4, 10, 25):
        so.coupled(bo, steps)
        sg.coupled(bg, steps)
        announce("after a call of", steps, "steps")
        assert max(abs(bo.cx - bg.cx), abs(bo.cy - bg.cy), abs(bo.cz - bg.cz)) <= 1e-4
        tag, tag_prev = sg.get("tag"), sg.get("tagPrev")
        assert np.array_equal(tag, so.tag)
        assert np.array_equal(tag_prev, so.tagPrev)
        fresh = np.flatnonzero(inner & (so.tag == TAG_FLUID) & (so.tagPrev == TAG_SOLID))
        dead = np.flatnonzero(inner & (so.tag == TAG_SOLID) & (so.tagPrev == TAG_FLUID))
        assert np.array_equal(np.asarray(sg.fresh_cells(), np.int64), fresh)
        assert np.array_equal(np.asarray(sg.dead_cells(), np.int64), dead)
        flips[0] += fresh.size
        flips[1] += dead.size
        mask = interior_fluid(so.tag, N, N, N)
        for k in ("phi", "rho", "ux", "uy", "uz"):
            e = rel_err(sg.get(k), getattr(so, k), mask)
            assert e <= 1e-5, (steps, k, e)
    # The sphere moved across cells: both lists were exercised.
    assert flips[0] > 0 and flips[1] > 0, flips
    sg.close()


@pytest.mark.parametrize("n", [256, 384])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_pair_kernel_equals_one_cell_kernel_fullsize(storage, n):
    """At 384^3 the 16-bit f buffer spans about 3 GB: the pair kernel's 32-bit
    buffer offsets then pass 2^31 bytes."""
    announce(storage, "pair vs one-cell stream kernel,", f"{n}^3 config-1 scene, 9 steps")
    out = []
    for lanes in (1, 2):
        sim, b = gpu_scene(storage, n)
        sim.set_stream_lanes(lanes)
        sim.coupled(b, 5)
        sim.coupled(b, 4)
        fields = {k: sim.get(k) for k in ("phi", "phiPrev", "rho", "ux", "uy", "uz", "tag",
                                          "tagPrev")}
        fluid = interior_fluid(fields["tag"], n, n, n)
        fields["f"] = pop_fluid(sim.get("f"), 27, fluid)
        fields["h"] = pop_fluid(sim.get("h"), 7, fluid)
        out.append((fields, b.as_list()))
        sim.close()
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k].view(np.uint8), out[1][0][k].view(np.uint8)), k
    assert out[0][1] == out[1][1]


# LBM_FULLSIZE_PREC_STEPS (a multiple of 50) lengthens the run, e.g. to the
# 2000 steps of the precision target.
PREC_STEPS = int(os.environ.get("LBM_FULLSIZE_PREC_STEPS", "200"))
PREC_SAMPLE = 50


@pytest.fixture(scope="module")
def fp64_256cube():
    announce("fp64 oracle (OpenMP build), 256^3 config-1 scene,", PREC_STEPS, "steps")
    o, b = oracle_scene("64_omp")
    traj = []
    for k in range(PREC_STEPS // PREC_SAMPLE):
        o.coupled(b, PREC_SAMPLE)
        traj.append((b.cx, b.cy, b.cz))
        announce("fp64 oracle step", (k + 1) * PREC_SAMPLE)
    return dict(phi=o.phi.copy(), ux=o.ux.copy(), uy=o.uy.copy(), uz=o.uz.copy(),
                tag=o.tag.copy(), traj=np.array(traj))


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_precision_vs_fp64_oracle_256cube(storage, fp64_256cube):
    announce("precision", storage, "256^3 config-1 scene,", PREC_STEPS, "steps")
    sim, b = gpu_scene(storage)
    traj = []
    for _ in range(PREC_STEPS // PREC_SAMPLE):
        sim.coupled(b, PREC_SAMPLE)
        traj.append((b.cx, b.cy, b.cz))
    g = dict(phi=sim.get("phi"), ux=sim.get("ux"), uy=sim.get("uy"), uz=sim.get("uz"),
             tag=sim.get("tag"), traj=np.array(traj))
    sim.close()
    e = errors(g, fp64_256cube)
    announce(storage, PREC_STEPS, "steps", e)
    assert np.isfinite(g["phi"]).all()
    assert e["phi_max"] <= PHI_MAX[storage], e
    assert e["sphere_max"] <= 0.05, e
