"""The CPU oracle against the golden vectors written by running the reference
simulator (tests/golden/make_goldens.mjs).  Every comparison is bitwise: the
oracle restates the reference's FP64 arithmetic and float32 stores exactly."""
import numpy as np
import pytest

import golden_io
import oracle_py as O
import scenes

FULL = ["f", "h", "phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev", "usx", "usy",
        "usz", "gx", "gy", "gz", "lap"]


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def mk_sphere(nx, ny, nz, cx, cy, cz, r, d):
    return O.sphere(nx, ny, nz, cx, cy, cz, r, d)


def check_trajectory(name, sim, sph, nsteps, fields):
    meta, A = golden_io.load(name)
    if sph is not None:
        assert sph.as_list() == meta["sphere0"]
    for k in fields:
        assert np.array_equal(bits(getattr(sim, k)), bits(A["step0/" + k])), ("step0", k)
    for t in range(1, nsteps + 1):
        sim.coupled(sph, 1)
        if sph is not None:
            assert sph.as_list() == meta["sphere"][t - 1], f"sphere differs at step {t}"
        if t in meta["snap_steps"]:
            for k in fields:
                assert np.array_equal(bits(getattr(sim, k)), bits(A[f"step{t}/" + k])), (t, k)


def test_lattice_tables():
    meta, A = golden_io.load("lattice_tables")
    import ctypes as C
    L = O.lib()[0]
    ex = (C.c_int8 * 27)(); ey = (C.c_int8 * 27)(); ez = (C.c_int8 * 27)()
    w = (C.c_float * 27)(); opp = (C.c_int8 * 27)()
    ex7 = (C.c_int8 * 7)(); ey7 = (C.c_int8 * 7)(); ez7 = (C.c_int8 * 7)()
    w7 = (C.c_float * 7)(); opp7 = (C.c_int8 * 7)()
    L.oracle_tables(ex, ey, ez, w, opp, ex7, ey7, ez7, w7, opp7)
    assert np.array_equal(np.array(w[:], np.float32), A["W27"])
    assert np.array_equal(np.array(w7[:], np.float32), A["W7"])
    for nm, arr in (("EX27", ex), ("EY27", ey), ("EZ27", ez), ("OPP27", opp), ("EX7", ex7),
                    ("EY7", ey7), ("EZ7", ez7), ("OPP7", opp7)):
        assert list(arr) == list(A[nm]), nm
    # float32-rounded weights, as printed by the reference (lattices.js:22).
    assert float(A["W27"][0]) == 0.29629629850387573
    assert float(A["W27"][1]) == 0.07407407462596893
    assert float(A["W27"][7]) == 0.018518518656492233
    assert float(A["W27"][19]) == 0.004629629664123058


def test_smoke12_trajectory():
    s = O.OSim(12, 12, 12)
    b = scenes.smoke12(s, mk_sphere)
    check_trajectory("smoke12", s, b, 80, FULL)


def test_config0_trajectory():
    s = O.OSim(40, 28, 40)
    b = scenes.config0(s, mk_sphere)
    check_trajectory("config0", s, b, 200, ["phi", "rho", "ux", "uy", "uz", "tag"])


def test_config0_fresh_dead_lists():
    meta, _ = golden_io.load("config0")
    s = O.OSim(40, 28, 40)
    b = scenes.config0(s, mk_sphere)
    for t in range(1, 201):
        s.coupled(b, 1)
        fresh = np.nonzero((s.tag == 0) & (s.tagPrev == 1))[0]
        dead = np.nonzero((s.tag == 1) & (s.tagPrev == 0))[0]
        assert list(fresh) == meta["fresh"][t - 1], t
        assert list(dead) == meta["dead"][t - 1], t


def test_dam_preset_host_writes():
    meta, _ = golden_io.load("dam16")
    nx, ny, nz = 16, 12, 14
    s = O.OSim(nx, ny, nz)
    s.set_params(**meta["params"])
    s.init_flat(ny + 1)
    for i, p in scenes.dam_fields(nx, ny, nz):
        s.phi[i] = p
        s.phiPrev[i] = p
        s.h[i * 7:(i + 1) * 7] = np.array(scenes.W7) * p
    s.seed_hydrostatic()
    b = O.sphere(nx, ny, nz, nx * 0.5, ny * 0.2, nz * 0.5, 4, 1.6)
    b.cx, b.cy, b.cz = nx * 0.7, ny * 0.6, nz * 0.5
    b.r = 3.5
    O.lib()[0].oracle_sphere_recompute_mass(b)
    b.density = 0.5
    O.lib()[0].oracle_sphere_recompute_mass(b)
    check_trajectory("dam16", s, b, 40, ["f", "h", "phi", "phiPrev", "rho", "ux", "uy", "uz",
                                         "tag", "tagPrev"])


def test_drag_spring_trajectory():
    s = O.OSim(20, 16, 20)
    b = scenes.drag20(s, mk_sphere)
    check_trajectory("drag20", s, b, 60, ["phi", "rho", "ux", "uy", "uz", "tag"])


def test_no_solid_profile():
    meta, A = golden_io.load("profile8")
    s = O.OSim(8, 12, 8)
    s.init_flat(6)
    s.set_params(tau=0.7, tauPhi=0.7, sigma=0.012, gravity=0.0012, rhoL=1.0, rhoG=0.1, W=4.0)
    for k in ("f", "h", "phi", "rho"):
        assert np.array_equal(bits(getattr(s, k)), bits(A["step0/" + k]))
    s.coupled(None, 60)
    for k in ("f", "h", "phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev"):
        assert np.array_equal(bits(getattr(s, k)), bits(A["step60/" + k])), k


def load_pass_state(sim, A, prefix="in"):
    for k in ["f", "f2", "h", "h2", "rho", "ux", "uy", "uz", "phi", "phiPrev", "usx", "usy",
              "usz", "gx", "gy", "gz", "lap", "bodyFx", "bodyFy", "bodyFz", "tag", "tagPrev"]:
        getattr(sim, k)[:] = A[f"{prefix}/{k}"]


@pytest.mark.parametrize("upto", range(7))
def test_single_passes(upto):
    """Each pass of step 11 of the kicked smoke scene, from the reference's own
    input state, against the reference's output arrays."""
    meta, A = golden_io.load("passes12")
    s = O.OSim(12, 12, 12)
    s.set_params(**meta["params"])
    b = O.OSphere()
    b.load_list(meta["sphere_in"])
    load_pass_state(s, A)
    fns = [lambda: s.retag(b), s.fresh, s.gradients, s.collide_phase, s.stream_phase,
           s.collide_hydro, lambda: s.stream_hydro_macro(b)]
    names = meta["passes"]
    # Replay earlier passes from the reference's recorded outputs, then run one.
    for j in range(upto):
        for k in meta["pass_outputs"][names[j]]:
            getattr(s, k)[:] = A[f"{names[j]}/{k}"]
    fns[upto]()
    for k in meta["pass_outputs"][names[upto]]:
        assert np.array_equal(bits(getattr(s, k)), bits(A[f"{names[upto]}/{k}"])), (names[upto], k)
    if names[upto] == "streamHydroAndMacro":
        assert b.as_list() == meta["sphere_after_stream"]


def test_absorb_step_and_pack():
    meta, A = golden_io.load("passes12")
    s = O.OSim(12, 12, 12)
    s.set_params(**meta["params"])
    b = O.OSphere()
    b.load_list(meta["sphere_in"])
    load_pass_state(s, A)
    s.step(b)
    assert b.as_list() == meta["sphere_after_stream"]
    s.absorb_dead(b)
    assert b.as_list() == meta["sphere_after_absorb"]
    O.sphere_step(b, 1.0, s.params()["gravity"])
    assert b.as_list() == meta["sphere_after_step"]
    assert np.array_equal(s.pack(), A["packVolume"])
