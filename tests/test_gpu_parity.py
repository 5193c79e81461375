"""HIP library against the reference's golden vectors and the CPU oracle.

Per-pass path: bitwise equal values to the reference for every pass, trajectory
and sphere state.  Fused path (lbm_step): fields equal to the per-pass path; the
sphere can differ in the last bits because link momentum is summed per block in
a fixed parallel order, so trajectory checks there use the stated tolerances:
phi/rho/u within 1e-5 of the field maximum, sphere centre within 1e-4 cells."""
import numpy as np
import pytest

import golden_io
import oracle_py as O
import scenes
from gpu_util import (MACROS, announce, gpu_sphere, interior_fluid, oracle_sphere, pop_fluid,
                      rel_err, same_values)
from lbm_py import lbm

pytestmark = pytest.mark.gpu

ALL_IN = ["f", "f2", "h", "h2", "rho", "ux", "uy", "uz", "phi", "phiPrev", "usx", "usy", "usz",
          "gx", "gy", "gz", "lap", "bodyFx", "bodyFy", "bodyFz", "tag", "tagPrev"]


def load_state(sim, A, prefix="in"):
    for k in ALL_IN:
        sim.set(k, A[f"{prefix}/{k}"])


def test_per_pass_kernels_match_reference():
    meta, A = golden_io.load("passes12")
    names = meta["passes"]
    for upto in range(7):
        announce("passes12", names[upto])
        sim = lbm.Sim(12, 12, 12)
        sim.set_params(**meta["params"])
        b = lbm.Sphere()
        b.load_list(meta["sphere_in"])
        load_state(sim, A)
        for j in range(upto):
            for k in meta["pass_outputs"][names[j]]:
                sim.set(k, A[f"{names[j]}/{k}"])
        if names[upto] == "handleFreshCells":
            # the fresh list comes from the retag; run it from the reference's
            # pre-retag state, then restore the reference's retag outputs.
            sim2 = lbm.Sim(12, 12, 12)
            sim2.set_params(**meta["params"])
            load_state(sim2, A)
            sim2.retag(b)
            fresh = sim2.fresh_cells()
            load_state(sim, A)
            sim.retag(b)
            for k in meta["pass_outputs"]["retagCells"]:
                assert same_values(sim.get(k), A[f"retagCells/{k}"]), k
            assert list(fresh) == list(sim.fresh_cells())
        fns = [lambda: sim.retag(b), sim.fresh, sim.gradients, sim.collide_phase, sim.stream_phase,
               sim.collide_hydro, lambda: sim.stream_hydro_macro(b)]
        fns[upto]()
        for k in meta["pass_outputs"][names[upto]]:
            assert same_values(sim.get(k), A[f"{names[upto]}/{k}"]), (names[upto], k)
        if names[upto] == "streamHydroAndMacro":
            assert b.as_list() == meta["sphere_after_stream"]
        sim.close()


def test_per_pass_absorb_and_pack():
    meta, A = golden_io.load("passes12")
    announce("passes12 full step + absorb + pack")
    sim = lbm.Sim(12, 12, 12)
    sim.set_params(**meta["params"])
    b = lbm.Sphere()
    b.load_list(meta["sphere_in"])
    load_state(sim, A)
    sim.step(b)
    assert b.as_list() == meta["sphere_after_stream"]
    sim.absorb_dead(b)
    assert b.as_list() == meta["sphere_after_absorb"]
    lbm.sphere_step(b, 1.0, meta["params"]["gravity"])
    assert b.as_list() == meta["sphere_after_step"]
    assert np.array_equal(sim.pack(), A["packVolume"])


def host_replay_step(sim, b):
    """Sim.step + absorbDeadCells with the impulses replayed on the host, one
    applyImpulseAtCell call per link in the reference's order
    (src/sim.js:509-543, src/solid.js:87-101)."""
    L = lbm.lib()
    sim.retag(b)
    sim.fresh()
    sim.gradients()
    sim.collide_phase()
    sim.stream_phase()
    sim.collide_hydro()
    links = sim.stream_hydro_macro_links()
    for (x, y, z, k, px, py, pz) in links:
        L.lbm_sphere_apply_impulse_at_cell(b, x, y, z, px, py, pz)
    p = sim.params()
    phi, ux, uy, uz = (sim.get(k) for k in ("phi", "ux", "uy", "uz"))
    dead = sim.dead_cells()
    for i in dead:
        i = int(i)
        rho = p["rhoG"] + float(phi[i]) * (p["rhoL"] - p["rhoG"])
        x, y, z = i % sim.nx, (i // sim.nx) % sim.ny, i // (sim.nx * sim.ny)
        L.lbm_sphere_apply_impulse_at_cell(b, x, y, z, rho * float(ux[i]), rho * float(uy[i]),
                                           rho * float(uz[i]))
    return len(links), len(dead)


@pytest.mark.parametrize("case", [
    # ~7.4k links: sorted in LDS
    ((48, 40, 36), (24.3, 18.2, 17.4, 8.0), (1, 8192)),
    # ~25-30k links, past the LDS sort: rocprim's radix sort orders the list.
    # The sphere starts clipped at the x = 0 face (the integrator then moves
    # it inside the walls, src/solid.js:198-205).
    ((64, 48, 56), (12.3, 22.7, 27.1, 16.0), (8193, 10 ** 9)),
    # a sphere of one or two cells: lists of a few links, or none
    ((20, 18, 22), (9.5, 8.5, 10.5, 0.8), (0, 64)),
], ids=["lds-sort", "rocprim-sort", "tiny"])
def test_per_pass_device_impulses_equal_host_replay(case):
    """lbm_sim_step and lbm_absorb_dead_cells apply the step's impulses to the
    sphere on the device; the sphere must equal the host replay of the same
    ordered impulses bit for bit, for a large, fast sphere (link lists of many
    1024-record chunks)."""
    (nx, ny, nz), (cx, cy, cz, r), (lo, hi) = case
    announce("device impulses vs host replay", (nx, ny, nz), r)

    def setup():
        sim = lbm.Sim(nx, ny, nz)
        sim.set_params(tau=0.6, tauPhi=0.6, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
        sim.init_flat(ny // 2)
        b = lbm.make_sphere(nx, ny, nz, cx, cy, cz, r, 1.8)
        b.vx, b.vy, b.vz = 0.35, -0.2, 0.15
        b.wx, b.wy, b.wz = 0.02, -0.03, 0.01
        return sim, b

    s1, b1 = setup()
    s2, b2 = setup()
    counts = []
    for t in range(6):
        s1.step(b1)
        s1.absorb_dead(b1)
        nl, nd = host_replay_step(s2, b2)
        counts.append(nl)
        assert b1.as_list() == b2.as_list(), f"sphere differs at step {t}"
        for k in ("phi", "rho", "ux", "uy", "uz", "tag"):
            assert same_values(s1.get(k), s2.get(k)), (t, k)
        lbm.sphere_step(b1, 1.0, 1e-4)
        lbm.sphere_step(b2, 1.0, 1e-4)
    assert lo <= min(counts) and max(counts) <= hi, counts
    s1.close()
    s2.close()


def run_golden(name, sim, b, nsteps, fields, fused, sphere_exact=True, field_tol=None):
    """Steps sim one call per step against the reference's recorded run: the
    sphere every step, the fresh and dead cell lists every step (when recorded),
    and the fields at the recorded snapshot steps."""
    meta, A = golden_io.load(name)
    if b is not None:
        assert b.as_list() == meta["sphere0"]
    worst = {}
    for t in range(1, nsteps + 1):
        if fused:
            sim.coupled(b, 1)
        else:
            sim.coupled_passes(b, 1)
        if b is not None:
            ref = meta["sphere"][t - 1]
            if sphere_exact:
                assert b.as_list() == ref, f"sphere differs at step {t}"
            else:
                d = max(abs(b.cx - ref[3]), abs(b.cy - ref[4]), abs(b.cz - ref[5]))
                assert d <= 1e-4, (t, d)
        if "fresh" in meta:
            # The reference's fresh and dead cells of this step's retag
            # (src/sim.js:246-256, src/solid.js:87-101), in visiting order.
            assert list(sim.fresh_cells()) == meta["fresh"][t - 1], ("fresh", t)
            assert list(sim.dead_cells()) == meta["dead"][t - 1], ("dead", t)
        if t in meta["snap_steps"]:
            tag = sim.get("tag")
            if "tag" in fields:
                assert np.array_equal(tag, A[f"step{t}/tag"]), t
            mask = interior_fluid(A[f"step{t}/tag"], sim.nx, sim.ny, sim.nz)
            for k in fields:
                if k in ("tag", "tagPrev"):
                    assert np.array_equal(sim.get(k), A[f"step{t}/{k}"]), (t, k)
                    continue
                got = sim.get(k)
                exp = A[f"step{t}/{k}"]
                if field_tol is None:
                    if k in ("f", "h"):
                        q = 27 if k == "f" else 7
                        assert same_values(pop_fluid(got, q, mask), pop_fluid(exp, q, mask)), (t, k)
                    else:
                        assert same_values(got, exp), (t, k)
                else:
                    e = rel_err(got, exp, mask if k not in ("f", "h") else None)
                    worst[k] = max(worst.get(k, 0), e)
                    assert e <= field_tol, (t, k, e)
    return worst


FULL = ["f", "h", "phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev", "usx", "usy", "usz",
        "gx", "gy", "gz", "lap"]


def test_per_pass_smoke12_trajectory():
    announce("per-pass smoke12 80 steps")
    sim = lbm.Sim(12, 12, 12)
    b = scenes.smoke12(sim, gpu_sphere)
    run_golden("smoke12", sim, b, 80, FULL, fused=False)


def test_per_pass_config0_trajectory():
    announce("per-pass config0 200 steps")
    sim = lbm.Sim(40, 28, 40)
    b = scenes.config0(sim, gpu_sphere)
    meta, _ = golden_io.load("config0")
    run_golden("config0", sim, b, 200, ["phi", "rho", "ux", "uy", "uz", "tag"], fused=False)


def test_fused_config0_trajectory():
    """The fast path on config 0: 200 coupled steps in one lbm_step call per step
    and in a single call, against the reference."""
    announce("fused config0 200 steps, one call per step")
    sim = lbm.Sim(40, 28, 40)
    b = scenes.config0(sim, gpu_sphere)
    worst = run_golden("config0", sim, b, 200, ["phi", "rho", "ux", "uy", "uz", "tag"], fused=True,
                       sphere_exact=False, field_tol=1e-5)
    announce("worst field error", worst)
    announce("fused config0 200 steps, one call")
    meta, A = golden_io.load("config0")
    sim2 = lbm.Sim(40, 28, 40)
    b2 = scenes.config0(sim2, gpu_sphere)
    sim2.coupled(b2, 200)
    ref = meta["sphere"][199]
    assert max(abs(b2.cx - ref[3]), abs(b2.cy - ref[4]), abs(b2.cz - ref[5])) <= 1e-4
    mask = interior_fluid(A["step200/tag"], 40, 28, 40)
    assert np.array_equal(sim2.get("tag"), A["step200/tag"])
    for k in ("phi", "rho", "ux", "uy", "uz"):
        assert rel_err(sim2.get(k), A[f"step200/{k}"], mask) <= 1e-5, k


def test_fused_smoke12_matches_reference():
    announce("fused smoke12 80 steps in one call")
    meta, A = golden_io.load("smoke12")
    sim = lbm.Sim(12, 12, 12)
    b = scenes.smoke12(sim, gpu_sphere)
    sim.coupled(b, 80)
    ref = meta["sphere"][79]
    assert max(abs(b.cx - ref[3]), abs(b.cy - ref[4]), abs(b.cz - ref[5])) <= 1e-4
    mask = interior_fluid(A["step80/tag"], 12, 12, 12)
    for k in ("phi", "rho", "ux", "uy", "uz"):
        assert rel_err(sim.get(k), A[f"step80/{k}"], mask) <= 1e-5, k
    assert np.array_equal(sim.get("tag"), A["step80/tag"])
    assert np.array_equal(sim.get("tagPrev"), A["step80/tagPrev"])


def test_fused_no_solid_is_exact():
    """Without a body the fused path has no summation-order freedom: results are
    the reference's bit for bit (test/profile.mjs scene)."""
    announce("fused profile8 60 steps, no solid")
    meta, A = golden_io.load("profile8")
    sim = lbm.Sim(8, 12, 8)
    sim.init_flat(6)
    sim.set_params(tau=0.7, tauPhi=0.7, sigma=0.012, gravity=0.0012, rhoL=1.0, rhoG=0.1, W=4.0)
    sim.coupled(None, 60)
    mask = interior_fluid(A["step60/tag"], 8, 12, 8)
    for k in ("phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev"):
        assert same_values(sim.get(k), A[f"step60/{k}"]), k
    for k, q in (("f", 27), ("h", 7)):
        assert same_values(pop_fluid(sim.get(k), q, mask), pop_fluid(A[f"step60/{k}"], q, mask)), k


def test_fused_dam_preset_host_writes():
    announce("fused dam16 40 steps")
    meta, A = golden_io.load("dam16")
    nx, ny, nz = 16, 12, 14
    sim = lbm.Sim(nx, ny, nz)
    sim.set_params(**meta["params"])
    sim.init_flat(ny + 1)
    phi = sim.get("phi")
    phiPrev = sim.get("phiPrev")
    h = sim.get("h")
    for i, p in scenes.dam_fields(nx, ny, nz):
        phi[i] = p
        phiPrev[i] = p
        h[i * 7:(i + 1) * 7] = np.array(scenes.W7) * p
    sim.set("phi", phi)
    sim.set("phiPrev", phiPrev)
    sim.set("h", h)
    sim.seed_hydrostatic()
    b = lbm.make_sphere(nx, ny, nz, nx * 0.5, ny * 0.2, nz * 0.5, 4, 1.6)
    b.cx, b.cy, b.cz = nx * 0.7, ny * 0.6, nz * 0.5
    b.r = 3.5
    lbm.recompute_mass(b)
    b.density = 0.5
    lbm.recompute_mass(b)
    run_golden("dam16", sim, b, 40, ["phi", "rho", "ux", "uy", "uz", "tag"], fused=True,
               sphere_exact=False, field_tol=1e-5)


def test_fused_drag_spring():
    announce("fused drag20 60 steps")
    sim = lbm.Sim(20, 16, 20)
    b = scenes.drag20(sim, gpu_sphere)
    run_golden("drag20", sim, b, 60, ["phi", "rho", "ux", "uy", "uz", "tag"], fused=True,
               sphere_exact=False, field_tol=1e-5)


def kicked_pair(n=24, steps=30):
    """A moving sphere with fresh and dead cells every step, run through the
    per-pass path and the fused path from the same state."""
    def setup():
        sim = lbm.Sim(n, n, n)
        sim.set_params(tau=0.7, tauPhi=0.62, sigma=0.01, gravity=2e-4, rhoL=1.0, rhoG=0.1, W=4.0)
        sim.init_flat(n * 0.55)
        b = lbm.make_sphere(n, n, n, n * 0.45, n * 0.35, n * 0.55, 4.3, 2.5)
        b.vx, b.vy, b.vz = 0.06, 0.03, -0.04
        b.wx, b.wy, b.wz = 0.01, -0.02, 0.015
        return sim, b
    return setup


def test_fused_equals_per_pass_moving_sphere():
    announce("fused vs per-pass, 24^3 kicked sphere, 30 steps")
    setup = kicked_pair()
    s1, b1 = setup()
    s2, b2 = setup()
    for t in range(30):
        s1.coupled_passes(b1, 1)
    s2.coupled(b2, 30)
    assert abs(b1.cx - b2.cx) <= 1e-9 and abs(b1.cy - b2.cy) <= 1e-9 and abs(b1.cz - b2.cz) <= 1e-9
    tag = s1.get("tag")
    assert np.array_equal(tag, s2.get("tag"))
    assert np.array_equal(s1.get("tagPrev"), s2.get("tagPrev"))
    assert np.array_equal(s1.fresh_cells(), s2.fresh_cells())
    assert np.array_equal(s1.dead_cells(), s2.dead_cells())
    mask = interior_fluid(tag, 24, 24, 24)
    for k in ("phi", "phiPrev", "rho", "ux", "uy", "uz"):
        assert rel_err(s2.get(k), s1.get(k), mask) <= 1e-9, k


def test_fused_equals_per_pass_single_steps_exact():
    """With the sphere held still (no force feedback through the integrator) the
    two paths are equal bit for bit, including populations and gradients."""
    announce("fused vs per-pass, held sphere, exact")
    setup = kicked_pair()
    s1, b1 = setup()
    s2, b2 = setup()
    for t in range(12):
        c1 = b1.copy()
        s1.coupled_passes(b1, 1)
        s2.coupled(c1, 1)
        # continue both from the per-pass sphere so the trajectories coincide
        b2 = b1.copy()
    tag = s1.get("tag")
    mask = interior_fluid(tag, 24, 24, 24)
    for k in ("phi", "phiPrev", "rho", "ux", "uy", "uz", "tag", "tagPrev", "gx", "gy", "gz", "lap",
              "usx", "usy", "usz"):
        assert same_values(s2.get(k), s1.get(k)), k
    for k, q in (("f", 27), ("h", 7), ("f2", 27), ("h2", 7)):
        assert same_values(pop_fluid(s2.get(k), q, mask), pop_fluid(s1.get(k), q, mask)), k


def test_fused_body_force_first_step():
    announce("fused vs per-pass with an external body force")
    setup = kicked_pair()
    s1, b1 = setup()
    s2, b2 = setup()
    rng = np.random.default_rng(3)
    bf = [(rng.standard_normal(s1.n) * 1e-4).astype(np.float32) for _ in range(3)]
    for s in (s1, s2):
        s.set("bodyFx", bf[0])
        s.set("bodyFy", bf[1])
        s.set("bodyFz", bf[2])
    s1.coupled_passes(b1, 3)
    s2.coupled(b2, 3)
    assert np.all(s2.get("bodyFx") == 0) and np.all(s1.get("bodyFy") == 0)
    mask = interior_fluid(s1.get("tag"), 24, 24, 24)
    for k in ("phi", "rho", "ux", "uy", "uz"):
        assert rel_err(s2.get(k), s1.get(k), mask) <= 1e-9, k


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_fused_is_deterministic(storage):
    """Two runs give the same bits everywhere, populations of wall and solid
    cells included (with 16-bit storage those are never written by the fused
    step and read back as zero deviations), whatever the device memory held."""
    announce("fused determinism, 32^3, 40 steps twice", storage)
    outs = []
    for rep in range(2):
        sim = lbm.Sim(32, 32, 32)
        sim.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
        sim.init_flat(18)
        sim.set_storage(storage)
        b = lbm.make_sphere(32, 32, 32, 15.3, 10.2, 16.7, 5.5, 3.0)
        b.vy = 0.05
        sim.coupled(b, 40)
        sim.coupled(b, 3)
        outs.append((b.as_list(), [sim.get(k) for k in ("phi", "rho", "ux", "f", "h")]))
        sim.close()
    assert outs[0][0] == outs[1][0]
    for a, c in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_fused_vs_oracle_larger_grid():
    """64x48x56 with a falling sphere: fused path against the oracle for 25 steps."""
    announce("fused vs oracle 64x48x56, 25 steps")
    nx, ny, nz = 64, 48, 56

    def setup(sim, mk):
        sim.set_params(tau=0.6, tauPhi=0.58, sigma=0.01, gravity=1e-4, rhoL=1.0, rhoG=0.1, W=4.0)
        sim.init_flat(ny * 0.6)
        b = mk(nx, ny, nz, 31.7, 14.2, 27.1, 7.5, 3.0)
        b.vx, b.vy, b.vz = 0.01, 0.05, -0.012
        return b

    so = O.OSim(nx, ny, nz)
    bo = setup(so, oracle_sphere)
    sg = lbm.Sim(nx, ny, nz)
    bg = setup(sg, gpu_sphere)
    so.coupled(bo, 25)
    sg.coupled(bg, 25)
    assert max(abs(bo.cx - bg.cx), abs(bo.cy - bg.cy), abs(bo.cz - bg.cz)) <= 1e-4
    assert np.array_equal(sg.get("tag"), so.tag)
    mask = interior_fluid(so.tag, nx, ny, nz)
    for k in ("phi", "rho", "ux", "uy", "uz"):
        e = rel_err(sg.get(k), getattr(so, k), mask)
        assert e <= 1e-5, (k, e)


def test_check_fields_flags_nan():
    """lbm_check_fields: finite and |u| <= 0.2 after steps; a NaN written into a
    fluid cell's phi is reported."""
    announce("check fields")
    sim = lbm.Sim(40, 28, 40)
    b = scenes.config0(sim, gpu_sphere)
    sim.coupled(b, 20)
    ok, um = sim.check_fields()
    assert ok and 0 < um <= 0.2 + 1e-7
    phi = sim.get("phi")
    tag = sim.get("tag")
    i = int(np.flatnonzero(tag == 0)[100])
    phi[i] = np.nan
    sim.set("phi", phi)
    ok, _ = sim.check_fields()
    assert not ok
    sim.close()
