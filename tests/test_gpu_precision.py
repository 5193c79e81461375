"""Compressed population storage against the FP64 oracle.

The fused step with LBM_STORAGE_F16 / LBM_STORAGE_BF16 stores populations as
16-bit deviations and computes the collide in FP32; LBM_STORAGE_F32 is the
reference-exact form (float32 storage, FP64 arithmetic).  Each is compared with
the oracle built with FP64 storage and arithmetic, on the two smallest scenes of
the seeded batched workload (64 scenes, axes 32..96, seed 2), after 2000 coupled
steps: max and RMS error of phi and u over cells fluid in both runs, and the
largest sphere-centre distance over the trajectory (sampled every 100 steps).
"""
import json
import os
import threading

import numpy as np
import pytest

import oracle_py as O
from gpu_util import announce
from lbm_py import lbm, scenegen

pytestmark = pytest.mark.gpu

STEPS = 2000
SAMPLE = 100
SCENE_IDS = None  # filled lazily: the two smallest scenes of config2(64, seed=2)
# Bounds on the max phi error after 2000 steps (the float32 form carries only
# the float32 storage rounding of the reference).
PHI_MAX = {"f32": 1e-4, "f16": 1e-3, "bf16": 1e-3}


def scenes():
    sc = scenegen.config2(64, seed=2)
    order = sorted(range(len(sc)), key=lambda i: np.prod(sc[i][0]))
    return [(i, sc[i]) for i in order[:2]]


def oracle_run(scene, out):
    (nx, ny, nz), params, yw, sph = scene
    o = O.OSim(nx, ny, nz, bits=64)
    o.set_params(**params)
    o.init_waterline_map(yw)
    b = O.sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"], bits=64)
    traj = []
    for _ in range(STEPS // SAMPLE):
        o.coupled(b, SAMPLE)
        traj.append((b.cx, b.cy, b.cz))
    out.update(phi=o.phi.copy(), ux=o.ux.copy(), uy=o.uy.copy(), uz=o.uz.copy(),
               tag=o.tag.copy(), traj=np.array(traj))


@pytest.fixture(scope="module")
def oracle_results():
    sc = scenes()
    res = [dict() for _ in sc]
    # ctypes releases the GIL inside the oracle, so the scenes run in parallel.
    th = [threading.Thread(target=oracle_run, args=(s, r)) for (_, s), r in zip(sc, res)]
    announce("fp64 oracle", [i for i, _ in sc], STEPS, "steps")
    for t in th:
        t.start()
    for t in th:
        t.join()
    return sc, res


def gpu_run(scene, storage):
    (nx, ny, nz), params, yw, sph = scene
    sim = lbm.Sim(nx, ny, nz)
    sim.set_params(**params)
    sim.init_waterline_map(yw)
    sim.set_storage(storage)
    b = lbm.make_sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
    traj = []
    for _ in range(STEPS // SAMPLE):
        sim.coupled(b, SAMPLE)
        traj.append((b.cx, b.cy, b.cz))
    out = dict(phi=sim.get("phi"), ux=sim.get("ux"), uy=sim.get("uy"), uz=sim.get("uz"),
               tag=sim.get("tag"), traj=np.array(traj), mass=sim.mass())
    sim.close()
    return out


def errors(g, o):
    both = (g["tag"] == 0) & (o["tag"] == 0)
    e = {}
    dphi = g["phi"][both].astype(np.float64) - o["phi"][both]
    e["phi_max"] = float(np.max(np.abs(dphi)))
    e["phi_rms"] = float(np.sqrt(np.mean(dphi ** 2)))
    du = np.sqrt(sum((g[k][both].astype(np.float64) - o[k][both]) ** 2 for k in ("ux", "uy", "uz")))
    umag = np.sqrt(sum(o[k][both] ** 2 for k in ("ux", "uy", "uz")))
    e["u_max"] = float(np.max(du))
    e["u_rms"] = float(np.sqrt(np.mean(du ** 2)))
    e["u_scale"] = float(np.max(umag))
    e["sphere_max"] = float(np.max(np.linalg.norm(g["traj"] - o["traj"], axis=1)))
    e["tag_mismatch"] = int(np.sum(g["tag"] != o["tag"]))
    return e


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_precision_vs_fp64_oracle(storage, oracle_results):
    sc, res = oracle_results
    report = []
    for (idx, scene), o in zip(sc, res):
        announce("precision", storage, "scene", idx, scene[0])
        g = gpu_run(scene, storage)
        e = errors(g, o)
        e.update(scene=idx, dims=list(scene[0]), storage=storage, steps=STEPS)
        report.append(e)
        print(json.dumps(e), flush=True)
        assert np.isfinite(g["phi"]).all()
        assert e["phi_max"] <= PHI_MAX[storage], e
        assert e["sphere_max"] <= 0.05, e
    path = os.environ.get("LBM_PRECISION_REPORT")
    if path:
        with open(path, "a") as fh:
            for e in report:
                fh.write(json.dumps(e) + "\n")


# The config-1 scene (perturbed waterline at 0.6 of the height, r = 24 sphere of
# density 3 at 0.2 of the height with a seeded velocity) at 128^3, run for the
# 2000 steps of the precision target against the FP64 oracle's OpenMP build.
# The 16-bit forms are held to the target, phi error <= 1e-3 (measured on
# MI355X: FP16 3.6e-4, BF16 4.1e-4).  The float32 form carries only the
# reference's own float32 storage rounding (measured 3.2e-4 here and 2.3e-4 on
# the 256^3 scene after 2000 steps): its bound is 5e-4, about 1.5 times that.
N128 = 128
PHI_MAX_128 = {"f32": 5e-4, "f16": 1e-3, "bf16": 1e-3}


def config1_128(kind, storage=None):
    params, yw, sph = scenegen.config1(N128, N128, N128, 1)
    if kind == "gpu":
        sim = lbm.Sim(N128, N128, N128)
        b = lbm.make_sphere(N128, N128, N128, sph["cx"], sph["cy"], sph["cz"], sph["r"],
                            sph["density"])
    else:
        sim = O.OSim(N128, N128, N128, bits="64_omp")
        b = O.sphere(N128, N128, N128, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"],
                     bits="64_omp")
    sim.set_params(**params)
    sim.init_waterline_map(yw)
    if storage:
        sim.set_storage(storage)
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    return sim, b


def run_sampled(sim, b):
    traj = []
    for k in range(STEPS // SAMPLE):
        sim.coupled(b, SAMPLE)
        traj.append((b.cx, b.cy, b.cz))
    return traj


@pytest.fixture(scope="module")
def fp64_128cube():
    announce("fp64 oracle (OpenMP build), 128^3 config-1 scene,", STEPS, "steps")
    o, b = config1_128("oracle")
    traj = run_sampled(o, b)
    return dict(phi=o.phi.copy(), ux=o.ux.copy(), uy=o.uy.copy(), uz=o.uz.copy(),
                tag=o.tag.copy(), traj=np.array(traj))


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_precision_2000_steps_config1_128cube(storage, fp64_128cube):
    announce("precision", storage, "128^3 config-1 scene,", STEPS, "steps")
    sim, b = config1_128("gpu", storage)
    traj = run_sampled(sim, b)
    g = dict(phi=sim.get("phi"), ux=sim.get("ux"), uy=sim.get("uy"), uz=sim.get("uz"),
             tag=sim.get("tag"), traj=np.array(traj))
    sim.close()
    e = errors(g, fp64_128cube)
    e.update(dims=[N128] * 3, storage=storage, steps=STEPS)
    print(json.dumps(e), flush=True)
    assert np.isfinite(g["phi"]).all()
    assert e["phi_max"] <= PHI_MAX_128[storage], e
    assert e["sphere_max"] <= 0.05, e
