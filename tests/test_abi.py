"""CPU-side checks of the C ABI: the library loads, exports every declared entry
point, and its host functions (lattice tables, RigidSphere) agree with the
oracle bit for bit.  No GPU compute is called."""
import ctypes as C
import math
import random
import re

import numpy as np
import pytest

import golden_io
import oracle_py as O
from lbm_py import HEADER, lbm


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|int64_t|char)\s*\*?\s*(lbm_\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = lbm.lib()
    names = declared_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(L, n), n
    assert L.lbm_version().decode().startswith("liblbm")


def test_lattice_tables_match_reference():
    meta, A = golden_io.load("lattice_tables")
    a = [(C.c_int32 * 27)() for _ in range(3)]
    w = (C.c_float * 27)()
    opp = (C.c_int32 * 27)()
    a7 = [(C.c_int32 * 7)() for _ in range(3)]
    w7 = (C.c_float * 7)()
    opp7 = (C.c_int32 * 7)()
    assert lbm.lib().lbm_lattice_tables(a[0], a[1], a[2], w, opp, a7[0], a7[1], a7[2], w7,
                                        opp7) == 0
    assert np.array_equal(np.array(w[:], np.float32), A["W27"])
    assert np.array_equal(np.array(w7[:], np.float32), A["W7"])
    for arr, key in zip(a + [opp] + a7 + [opp7],
                        ["EX27", "EY27", "EZ27", "OPP27", "EX7", "EY7", "EZ7", "OPP7"]):
        assert list(arr) == list(A[key]), key


def to_oracle(b):
    o = O.OSphere()
    o.load_list(b.as_list())
    return o


def test_sphere_defaults_match_reference_constructor():
    # new RigidSphere(40, 28, 40, {}) (src/solid.js:11-14, 24)
    b = lbm.make_sphere(40, 28, 40)
    assert (b.cx, b.cy, b.cz) == (20.0, 28 * 0.3, 20.0)
    assert b.r == min(40, 28, 40) * 0.10 and b.density == 1.4
    assert b.mass == b.density * ((4 / 3) * math.pi * b.r * b.r * b.r)


def test_sphere_functions_bitwise_vs_oracle():
    rng = random.Random(7)
    L = lbm.lib()
    for trial in range(400):
        b = lbm.make_sphere(40, 28, 40, rng.uniform(5, 35), rng.uniform(4, 24),
                            rng.uniform(5, 35), rng.uniform(1, 8), rng.uniform(0.3, 4))
        for k in ("vx", "vy", "vz"):
            setattr(b, k, rng.uniform(-0.1, 0.1))
        for k in ("wx", "wy", "wz"):
            setattr(b, k, rng.uniform(-0.05, 0.05))
        if trial % 3 == 0:
            b.dragging = 1.0
            b.dragTargetX, b.dragTargetY, b.dragTargetZ = (rng.uniform(0, 40), rng.uniform(0, 28),
                                                           rng.uniform(0, 40))
        o = to_oracle(b)
        x, y, z = rng.randrange(40), rng.randrange(28), rng.randrange(40)
        assert L.lbm_sphere_contains(C.byref(b), x, y, z) == O.lib()[0].oracle_sphere_contains(
            C.byref(o), x, y, z)
        out1 = (C.c_double * 3)()
        out2 = (C.c_double * 3)()
        L.lbm_sphere_surface_velocity(C.byref(b), x, y, z, out1)
        O.lib()[0].oracle_sphere_surface_velocity(C.byref(o), x, y, z, out2)
        assert list(out1) == list(out2)
        for _ in range(5):
            p = [rng.uniform(-1e-2, 1e-2) for _ in range(3)]
            L.lbm_sphere_apply_impulse_at_cell(C.byref(b), x, y, z, *p)
            O.lib()[0].oracle_sphere_apply_impulse_at_cell(C.byref(o), x, y, z, *p)
        # Large forces exercise the acceleration and speed caps; positions near
        # the walls exercise the clamp.
        if trial % 4 == 1:
            b.Fy *= 1e4
            o.Fy = b.Fy
        g = [1e-5, 0.0015, 0.0, float("nan")][trial % 4] if trial % 17 else 0.001
        lbm.sphere_step(b, 1.0, g)
        O.sphere_step(o, 1.0, g)
        la, lo = b.as_list(), o.as_list()
        for va, vo in zip(la, lo):
            assert (va == vo) or (math.isnan(va) and math.isnan(vo)), (trial, la, lo)


def test_sphere_step_golden_sequence():
    """The host integrator reproduces the reference's sphere trajectory when fed
    the reference's own per-step forces (config 0, 200 steps)."""
    meta, _ = golden_io.load("config0")
    b = lbm.Sphere()
    b.load_list(meta["sphere0"])
    for t in range(200):
        fl = meta["forces"][t][1]
        b.Fx, b.Fy, b.Fz, b.Tx, b.Ty, b.Tz = fl
        lbm.sphere_step(b, 1.0, meta["params"]["gravity"])
        assert b.as_list() == meta["sphere"][t], t


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = lbm.lib().lbm_create(8, 8, 8, C.byref(h))
    assert rc == -6 and not h.value
    assert b"no HIP device" in lbm.lib().lbm_last_error(None)
