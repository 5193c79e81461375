"""ctypes binding of liblbm (include/lbm.h).

Host-side plumbing for Python callers (tests, the benchmark): it loads
lib/liblbm.so and exposes the C entry points with the same method names as the
reference's Sim / RigidSphere (src/sim.js, src/solid.js).  All computation
happens in the HIP library; this module only moves arguments."""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liblbm.so")

SPHERE_FIELDS = ["nx", "ny", "nz", "cx", "cy", "cz", "r", "vx", "vy", "vz", "wx", "wy", "wz",
                 "density", "volume", "mass", "I", "dragging", "dragTargetX", "dragTargetY",
                 "dragTargetZ", "Fx", "Fy", "Fz", "Tx", "Ty", "Tz"]
PARAMS = ["tau", "tauPhi", "sigma", "gravity", "rhoL", "rhoG", "W", "mobility"]

FIELD_IDS = {name: i for i, name in enumerate(
    ["f", "f2", "h", "h2", "rho", "ux", "uy", "uz", "phi", "phiPrev", "tag", "tagPrev",
     "usx", "usy", "usz", "gx", "gy", "gz", "lap", "bodyFx", "bodyFy", "bodyFz"])}
_Q = {"f": 27, "f2": 27, "h": 7, "h2": 7}
_U8 = {"tag", "tagPrev"}


class LbmError(RuntimeError):
    pass


class Sphere(C.Structure):
    """struct lbm_sphere: the fields of RigidSphere (src/solid.js:9-37)."""
    _fields_ = [(k, C.c_double) for k in SPHERE_FIELDS]

    def as_list(self):
        return [getattr(self, k) for k in SPHERE_FIELDS]

    def load_list(self, vals):
        for k, v in zip(SPHERE_FIELDS, vals):
            setattr(self, k, float(v))

    def copy(self):
        s = Sphere()
        C.memmove(C.byref(s), C.byref(self), C.sizeof(Sphere))
        return s


class Params(C.Structure):
    _fields_ = [(k, C.c_double) for k in PARAMS]


class Link(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("k", C.c_int32),
                ("px", C.c_double), ("py", C.c_double), ("pz", C.c_double)]


class Request(C.Structure):
    """struct lbm_request."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("params", Params),
                ("waterline", C.POINTER(C.c_double)), ("waterline_y", C.c_double),
                ("sphere", Sphere), ("has_sphere", C.c_int), ("steps", C.c_int),
                ("frame_every", C.c_int), ("storage", C.c_int)]


class FrameInfo(C.Structure):
    """struct lbm_frame_info."""
    _fields_ = [("request", C.c_int64), ("status", C.c_int), ("index", C.c_int),
                ("step", C.c_int), ("last", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("bytes", C.c_int64), ("t_submit", C.c_double),
                ("t_ready", C.c_double), ("latency", C.c_double), ("sphere", Sphere)]


class ServerStats(C.Structure):
    """struct lbm_server_stats."""
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("rejected", C.c_int64), ("active", C.c_int), ("pending", C.c_int),
                ("max_active_seen", C.c_int), ("queued_frames", C.c_int),
                ("frames", C.c_int64), ("ticks", C.c_int64), ("lattice_updates", C.c_int64),
                ("busy_seconds", C.c_double), ("uptime_seconds", C.c_double),
                ("bytes_in_use", C.c_int64), ("budget", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    """lbm_tune_result: the launch configuration the autotuner chose."""
    _fields_ = [("cells_per_lane", C.c_int), ("graph_steps", C.c_int),
                ("ms_per_step", C.c_double), ("candidates", C.c_int), ("stream_split", C.c_int),
                ("march_rows", C.c_int), ("march_segments", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


STORAGE = {"f32": 0, "f16": 1, "bf16": 2}


def lib():
    """Loads liblbm; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbmError(f"{LIB_PATH} is missing: run build.py (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    SP = C.POINTER(Sphere)
    sig = {
        "lbm_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
        "lbm_destroy": (None, [P]),
        "lbm_last_error": (C.c_char_p, [P]),
        "lbm_dims": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "lbm_set_params": (C.c_int, [P, C.POINTER(Params)]),
        "lbm_get_params": (C.c_int, [P, C.POINTER(Params)]),
        "lbm_init_flat": (C.c_int, [P, C.c_double]),
        "lbm_init_waterline_map": (C.c_int, [P, C.POINTER(C.c_double)]),
        "lbm_seed_hydrostatic": (C.c_int, [P]),
        "lbm_retag_cells": (C.c_int, [P, SP]),
        "lbm_retag_cells_mask": (C.c_int, [P, C.POINTER(C.c_uint8), C.POINTER(C.c_float)]),
        "lbm_handle_fresh_cells": (C.c_int, [P]),
        "lbm_compute_phi_gradients": (C.c_int, [P]),
        "lbm_collide_phase": (C.c_int, [P]),
        "lbm_stream_phase": (C.c_int, [P]),
        "lbm_collide_hydro": (C.c_int, [P]),
        "lbm_stream_hydro_macro": (C.c_int, [P, SP]),
        "lbm_stream_hydro_macro_links": (C.c_int, [P, C.POINTER(C.c_int)]),
        "lbm_get_links": (C.c_int, [P, C.POINTER(Link), C.c_int]),
        "lbm_sim_step": (C.c_int, [P, SP]),
        "lbm_absorb_dead_cells": (C.c_int, [P, SP]),
        "lbm_step": (C.c_int, [P, SP, C.c_int]),
        "lbm_step_noalloc": (C.c_int, [P, SP, C.c_int]),
        "lbm_step_live": (C.c_int, [P, SP, C.c_int]),
        "lbm_set_stream_split": (C.c_int, [P, C.c_int]),
        "lbm_set_march_shape": (C.c_int, [P, C.c_int, C.c_int]),
        "lbm_batch_set_stream_split": (C.c_int, [P, C.c_int]),
        "lbm_field_count": (C.c_int64, [P, C.c_int]),
        "lbm_download": (C.c_int, [P, C.c_int, P]),
        "lbm_upload": (C.c_int, [P, C.c_int, P]),
        "lbm_pack_volume": (C.c_int, [P, C.POINTER(C.c_uint8)]),
        "lbm_set_storage": (C.c_int, [P, C.c_int]),
        "lbm_get_storage": (C.c_int, [P]),
        "lbm_mass": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "lbm_check_fields": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "lbm_fresh_cells": (C.c_int, [P, C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_int)]),
        "lbm_dead_cells": (C.c_int, [P, C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_int)]),
        "lbm_sphere_init": (None, [SP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double]),
        "lbm_sphere_recompute_mass": (None, [SP]),
        "lbm_sphere_contains": (C.c_int, [SP, C.c_int, C.c_int, C.c_int]),
        "lbm_sphere_surface_velocity": (None, [SP, C.c_int, C.c_int, C.c_int,
                                               C.POINTER(C.c_double)]),
        "lbm_sphere_apply_impulse_at_cell": (None, [SP, C.c_int, C.c_int, C.c_int, C.c_double,
                                                    C.c_double, C.c_double]),
        "lbm_sphere_step": (None, [SP, C.c_double, C.c_double]),
        "lbm_lattice_tables": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
        "lbm_stream": (P, [P]),
        "lbm_synchronize": (C.c_int, [P]),
        "lbm_set_timing": (C.c_int, [P, C.c_int]),
        "lbm_kernel_time": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]),
        "lbm_set_graph_steps": (C.c_int, [P, C.c_int]),
        "lbm_graph_counts": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "lbm_set_stream_lanes": (C.c_int, [P, C.c_int]),
        "lbm_version": (C.c_char_p, []),
        "lbm_probe_copy_bandwidth": (C.c_int, [C.c_int64, C.c_int, C.POINTER(C.c_double)]),
        "lbm_probe_pull_bandwidth": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.POINTER(C.c_double)]),
        "lbm_state_bytes": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "lbm_server_create": (C.c_int, [C.c_int64, C.c_int, C.c_int64, C.POINTER(P)]),
        "lbm_server_destroy": (None, [P]),
        "lbm_server_last_error": (C.c_char_p, [P]),
        "lbm_server_submit": (C.c_int, [P, C.POINTER(Request), C.POINTER(C.c_int64)]),
        "lbm_server_next_frame": (C.c_int, [P, C.c_int, C.POINTER(FrameInfo),
                                            C.POINTER(C.c_uint8), C.c_int64]),
        "lbm_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
        "lbm_server_clock": (C.c_double, [P]),
        "lbm_server_pause": (C.c_int, [P, C.c_int]),
        "lbm_batch_create": (C.c_int, [C.POINTER(P)]),
        "lbm_batch_destroy": (None, [P]),
        "lbm_batch_last_error": (C.c_char_p, [P]),
        "lbm_batch_add": (C.c_int, [P, P, SP]),
        "lbm_batch_remove": (C.c_int, [P, P]),
        "lbm_batch_size": (C.c_int, [P]),
        "lbm_batch_step": (C.c_int, [P, C.c_int]),
        "lbm_batch_set_timing": (C.c_int, [P, C.c_int]),
        "lbm_batch_kernel_time": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int64)]),
        "lbm_batch_set_graph_steps": (C.c_int, [P, C.c_int]),
        "lbm_batch_graph_counts": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "lbm_batch_set_stream_lanes": (C.c_int, [P, C.c_int]),
        "lbm_batch_autotune": (C.c_int, [P, C.c_int, C.POINTER(TuneResult)]),
        "lbm_autotune": (C.c_int, [P, SP, C.c_int, C.POINTER(TuneResult)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def make_sphere(nx, ny, nz, cx=float("nan"), cy=float("nan"), cz=float("nan"),
                r=float("nan"), density=float("nan")):
    """new RigidSphere(nx, ny, nz, {cx, cy, cz, r, density}); NaN = default."""
    b = Sphere()
    lib().lbm_sphere_init(C.byref(b), nx, ny, nz, cx, cy, cz, r, density)
    return b


def pull_bandwidth(nx, ny, nz, storage="f32", reps=10):
    """Measured read+write GB/s of the stream kernels' population pattern alone
    (34 shifted planes pulled and stored, no arithmetic) on an nx x ny x nz grid."""
    g = C.c_double(0)
    rc = lib().lbm_probe_pull_bandwidth(int(nx), int(ny), int(nz), STORAGE[storage], int(reps),
                                        C.byref(g))
    if rc != 0:
        raise LbmError(f"lbm_probe_pull_bandwidth = {rc}")
    return g.value


def copy_bandwidth(nbytes=1 << 31, reps=10):
    """Measured read+write GB/s of a plain device copy of nbytes."""
    g = C.c_double(0)
    rc = lib().lbm_probe_copy_bandwidth(int(nbytes), int(reps), C.byref(g))
    if rc != 0:
        raise LbmError(f"lbm_probe_copy_bandwidth = {rc}")
    return g.value


def sphere_step(b, dt, gravity):
    lib().lbm_sphere_step(C.byref(b), dt, gravity)


def recompute_mass(b):
    lib().lbm_sphere_recompute_mass(C.byref(b))


class Sim:
    """A device-resident simulation (struct lbm_sim)."""

    def __init__(self, nx, ny, nz):
        L = lib()
        h = C.c_void_p()
        rc = L.lbm_create(nx, ny, nz, C.byref(h))
        if rc != 0:
            raise LbmError(f"lbm_create({nx},{ny},{nz}) = {rc}: "
                           f"{L.lbm_last_error(None).decode()}")
        self.h = h
        self.L = L
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            raise LbmError(f"{what} = {rc}: {self.L.lbm_last_error(self.h).decode()}")

    # -- parameters
    def set_params(self, **kw):
        p = Params()
        self._ck(self.L.lbm_get_params(self.h, C.byref(p)), "lbm_get_params")
        for k, v in kw.items():
            if k not in PARAMS:
                raise KeyError(k)
            setattr(p, k, float(v))
        self._ck(self.L.lbm_set_params(self.h, C.byref(p)), "lbm_set_params")

    def params(self):
        p = Params()
        self._ck(self.L.lbm_get_params(self.h, C.byref(p)), "lbm_get_params")
        return {k: getattr(p, k) for k in PARAMS}

    # -- init
    def init_flat(self, y):
        self._ck(self.L.lbm_init_flat(self.h, float(y)), "lbm_init_flat")

    def init_waterline_map(self, yw):
        yw = np.ascontiguousarray(yw, dtype=np.float64)
        assert yw.size == self.nx * self.nz
        self._ck(self.L.lbm_init_waterline_map(self.h, yw.ctypes.data_as(C.POINTER(C.c_double))),
                 "lbm_init_waterline_map")

    def seed_hydrostatic(self):
        self._ck(self.L.lbm_seed_hydrostatic(self.h), "lbm_seed_hydrostatic")

    # -- passes (reference array semantics)
    @staticmethod
    def _sp(b):
        return C.byref(b) if b is not None else None

    def retag(self, b):
        self._ck(self.L.lbm_retag_cells(self.h, self._sp(b)), "lbm_retag_cells")

    def retag_mask(self, mask, us):
        mask = np.ascontiguousarray(mask, np.uint8)
        us = np.ascontiguousarray(us, np.float32)
        self._ck(self.L.lbm_retag_cells_mask(self.h, mask.ctypes.data_as(C.POINTER(C.c_uint8)),
                                             us.ctypes.data_as(C.POINTER(C.c_float))),
                 "lbm_retag_cells_mask")

    def fresh(self):
        self._ck(self.L.lbm_handle_fresh_cells(self.h), "lbm_handle_fresh_cells")

    def gradients(self):
        self._ck(self.L.lbm_compute_phi_gradients(self.h), "lbm_compute_phi_gradients")

    def collide_phase(self):
        self._ck(self.L.lbm_collide_phase(self.h), "lbm_collide_phase")

    def stream_phase(self):
        self._ck(self.L.lbm_stream_phase(self.h), "lbm_stream_phase")

    def collide_hydro(self):
        self._ck(self.L.lbm_collide_hydro(self.h), "lbm_collide_hydro")

    def stream_hydro_macro(self, b):
        self._ck(self.L.lbm_stream_hydro_macro(self.h, self._sp(b)), "lbm_stream_hydro_macro")

    def stream_hydro_macro_links(self):
        """streamHydroAndMacro for a generic solid: returns the ordered
        applyImpulseAtCell calls as (x, y, z, k, px, py, pz) tuples."""
        cnt = C.c_int(0)
        self._ck(self.L.lbm_stream_hydro_macro_links(self.h, C.byref(cnt)),
                 "lbm_stream_hydro_macro_links")
        buf = (Link * max(cnt.value, 1))()
        self._ck(self.L.lbm_get_links(self.h, buf, cnt.value), "lbm_get_links")
        return [(l.x, l.y, l.z, l.k, l.px, l.py, l.pz) for l in buf[:cnt.value]]

    def step(self, b):
        """Sim.step(solid): the seven passes."""
        self._ck(self.L.lbm_sim_step(self.h, self._sp(b)), "lbm_sim_step")

    def absorb_dead(self, b):
        self._ck(self.L.lbm_absorb_dead_cells(self.h, C.byref(b)), "lbm_absorb_dead_cells")

    def coupled_passes(self, b, nsteps):
        """nsteps of sim.step; absorbDeadCells; sphere step, pass by pass."""
        for _ in range(nsteps):
            self.step(b)
            if b is not None:
                self.absorb_dead(b)
                sphere_step(b, 1.0, self.params()["gravity"])

    def coupled(self, b, nsteps):
        """nsteps fused coupled steps (lbm_step)."""
        self._ck(self.L.lbm_step(self.h, self._sp(b), int(nsteps)), "lbm_step")

    def coupled_live(self, b, nsteps):
        """lbm_step_live: steps that end live (the next call continues without
        the per-call prologue); only phi, the tags, lists and sphere are the
        reference's until a coupled() call ends the run."""
        self._ck(self.L.lbm_step_live(self.h, self._sp(b), int(nsteps)), "lbm_step_live")

    def coupled_noalloc(self, b, nsteps):
        """lbm_step_noalloc: the same steps without device allocation."""
        self._ck(self.L.lbm_step_noalloc(self.h, self._sp(b), int(nsteps)), "lbm_step_noalloc")

    # -- fields
    def get(self, name):
        q = _Q.get(name, 1)
        dt = np.uint8 if name in _U8 else np.float32
        out = np.empty(self.n * q, dtype=dt)
        self._ck(self.L.lbm_download(self.h, FIELD_IDS[name], out.ctypes.data_as(C.c_void_p)),
                 f"lbm_download({name})")
        return out

    def set(self, name, arr):
        q = _Q.get(name, 1)
        dt = np.uint8 if name in _U8 else np.float32
        a = np.ascontiguousarray(arr, dtype=dt)
        assert a.size == self.n * q, (name, a.size)
        self._ck(self.L.lbm_upload(self.h, FIELD_IDS[name], a.ctypes.data_as(C.c_void_p)),
                 f"lbm_upload({name})")

    def __getattr__(self, name):
        if name in FIELD_IDS:
            return self.get(name)
        raise AttributeError(name)

    def set_storage(self, kind):
        """Population storage of the fused step: 'f32' (reference-exact), 'f16' or
        'bf16' (16-bit deviations, FP32 collide arithmetic)."""
        self._ck(self.L.lbm_set_storage(self.h, STORAGE[kind]), "lbm_set_storage")

    @property
    def storage(self):
        k = self.L.lbm_get_storage(self.h)
        return {v: n for n, v in STORAGE.items()}[k]

    def mass(self):
        """FP64 sums of phi and of rho over fluid cells."""
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.lbm_mass(self.h, C.byref(a), C.byref(b)), "lbm_mass")
        return a.value, b.value

    def check_fields(self):
        """(all finite, max |u|) over fluid cells."""
        ok, um = C.c_int(), C.c_double()
        self._ck(self.L.lbm_check_fields(self.h, C.byref(ok), C.byref(um)), "lbm_check_fields")
        return bool(ok.value), um.value

    def pack(self):
        out = np.empty(2 * self.n, np.uint8)
        self._ck(self.L.lbm_pack_volume(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8))),
                 "lbm_pack_volume")
        return out

    def _list(self, fn, what):
        cnt = C.c_int(0)
        self._ck(fn(self.h, None, 0, C.byref(cnt)), what)
        out = np.empty(max(cnt.value, 1), np.int32)
        self._ck(fn(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value, C.byref(cnt)),
                 what)
        return out[:cnt.value]

    def fresh_cells(self):
        return self._list(self.L.lbm_fresh_cells, "lbm_fresh_cells")

    def dead_cells(self):
        return self._list(self.L.lbm_dead_cells, "lbm_dead_cells")

    # -- timing
    def stream_handle(self):
        return self.L.lbm_stream(self.h)

    def synchronize(self):
        self._ck(self.L.lbm_synchronize(self.h), "lbm_synchronize")

    def set_timing(self, on):
        self._ck(self.L.lbm_set_timing(self.h, 1 if on else 0), "lbm_set_timing")

    def kernel_time(self, name):
        ms = C.c_double(0)
        cnt = C.c_int64(0)
        self._ck(self.L.lbm_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(cnt)),
                 "lbm_kernel_time")
        return ms.value, cnt.value

    def set_graph_steps(self, steps):
        """Steps per captured launch graph of coupled(); 0 launches kernels one by one."""
        self._ck(self.L.lbm_set_graph_steps(self.h, int(steps)), "lbm_set_graph_steps")

    def set_stream_split(self, mode):
        """lbm_set_stream_split: 1 bulk/band-row split, 0 one launch, 2 the row
        march (float32, rows of at most 256 cells)."""
        self._ck(self.L.lbm_set_stream_split(self.h, int(mode)), "lbm_set_stream_split")

    def set_march_shape(self, rows=2, segments=0):
        self._ck(self.L.lbm_set_march_shape(self.h, int(rows), int(segments)),
                 "lbm_set_march_shape")

    def set_stream_lanes(self, cells_per_lane):
        """Cells per lane (1 or 2) of the 16-bit stream kernel."""
        self._ck(self.L.lbm_set_stream_lanes(self.h, int(cells_per_lane)), "lbm_set_stream_lanes")

    def autotune(self, sphere=None, trial_steps=10):
        """Times the launch configurations over real steps (every one computes
        the same state), keeps the fastest; returns it as a dict."""
        r = TuneResult()
        self._ck(self.L.lbm_autotune(self.h, C.byref(sphere) if sphere is not None else None,
                                     int(trial_steps), C.byref(r)), "lbm_autotune")
        return r.as_dict()

    def graph_counts(self):
        """(captures, replays) of launch graphs since the handle was created."""
        a, b = C.c_int64(0), C.c_int64(0)
        self._ck(self.L.lbm_graph_counts(self.h, C.byref(a), C.byref(b)), "lbm_graph_counts")
        return a.value, b.value


class Batch:
    """Many simulations stepped by the same launches (struct lbm_batch)."""

    def __init__(self):
        L = lib()
        h = C.c_void_p()
        rc = L.lbm_batch_create(C.byref(h))
        if rc != 0:
            raise LbmError(f"lbm_batch_create = {rc}: {L.lbm_last_error(None).decode()}")
        self.h = h
        self.L = L
        self.members = []  # keeps the Sim and Sphere objects alive

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc < 0:
            raise LbmError(f"{what} = {rc}: {self.L.lbm_batch_last_error(self.h).decode()}")
        return rc

    def add(self, sim, sphere=None):
        idx = self._ck(self.L.lbm_batch_add(self.h, sim.h, C.byref(sphere) if sphere is not None
                                            else None), "lbm_batch_add")
        self.members.append((sim, sphere))
        return idx

    def remove(self, sim):
        self._ck(self.L.lbm_batch_remove(self.h, sim.h), "lbm_batch_remove")
        self.members = [m for m in self.members if m[0] is not sim]

    def step(self, nsteps):
        self._ck(self.L.lbm_batch_step(self.h, int(nsteps)), "lbm_batch_step")

    def set_timing(self, on):
        self._ck(self.L.lbm_batch_set_timing(self.h, 1 if on else 0), "lbm_batch_set_timing")

    def kernel_time(self, name):
        ms = C.c_double(0)
        cnt = C.c_int64(0)
        self._ck(self.L.lbm_batch_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(cnt)),
                 "lbm_batch_kernel_time")
        return ms.value, cnt.value

    def set_graph_steps(self, steps):
        self._ck(self.L.lbm_batch_set_graph_steps(self.h, int(steps)), "lbm_batch_set_graph_steps")

    def set_stream_split(self, on):
        """lbm_batch_set_stream_split: 1 split, 0 one launch, -1 by storage."""
        rc = self.L.lbm_batch_set_stream_split(self.h, int(on))
        if rc != 0:
            raise LbmError(f"lbm_batch_set_stream_split = {rc}")

    def set_stream_lanes(self, cells_per_lane):
        self._ck(self.L.lbm_batch_set_stream_lanes(self.h, int(cells_per_lane)),
                 "lbm_batch_set_stream_lanes")

    def autotune(self, trial_steps=10):
        """lbm_batch_autotune: returns the chosen configuration as a dict."""
        r = TuneResult()
        self._ck(self.L.lbm_batch_autotune(self.h, int(trial_steps), C.byref(r)),
                 "lbm_batch_autotune")
        return r.as_dict()

    def graph_counts(self):
        a, b = C.c_int64(0), C.c_int64(0)
        self._ck(self.L.lbm_batch_graph_counts(self.h, C.byref(a), C.byref(b)),
                 "lbm_batch_graph_counts")
        return a.value, b.value


_live_servers = weakref.WeakSet()


@atexit.register
def _close_servers():
    # The loop thread must stop before the process tears the HIP runtime down.
    for srv in list(_live_servers):
        srv.close()


class Server:
    """The serving loop (struct lbm_server): submit requests, read RG8 frames."""

    def __init__(self, hbm_budget, max_active=256, frame_queue_bytes=0):
        L = lib()
        h = C.c_void_p()
        rc = L.lbm_server_create(int(hbm_budget), int(max_active), int(frame_queue_bytes),
                                 C.byref(h))
        if rc != 0:
            raise LbmError(f"lbm_server_create = {rc}: {L.lbm_last_error(None).decode()}")
        self.h = h
        self.L = L
        self._buf = np.empty(0, np.uint8)
        _live_servers.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc < 0:
            raise LbmError(f"{what} = {rc}: {self.L.lbm_server_last_error(self.h).decode()}")
        return rc

    def submit(self, nx, ny, nz, params, steps, frame_every=10, waterline=None, waterline_y=0.0,
               sphere=None, storage="f32"):
        """Queues a scene; waterline is an [nz, nx] array of interface heights
        (lbm_init_waterline_map) or None for a flat interface at waterline_y."""
        r = Request()
        r.nx, r.ny, r.nz = nx, ny, nz
        for k in PARAMS:
            setattr(r.params, k, float(params.get(k, 0.02 if k == "mobility" else 0.0)))
        keep = None
        if waterline is not None:
            keep = np.ascontiguousarray(waterline, dtype=np.float64).ravel()
            assert keep.size == nx * nz
            r.waterline = keep.ctypes.data_as(C.POINTER(C.c_double))
        r.waterline_y = float(waterline_y)
        if sphere is not None:
            C.memmove(C.byref(r.sphere), C.byref(sphere), C.sizeof(Sphere))
            r.has_sphere = 1
        r.steps, r.frame_every, r.storage = int(steps), int(frame_every), STORAGE[storage]
        if self._buf.size < 2 * nx * ny * nz:
            self._buf = np.empty(2 * nx * ny * nz, np.uint8)
        rid = C.c_int64()
        self._ck(self.L.lbm_server_submit(self.h, C.byref(r), C.byref(rid)), "lbm_server_submit")
        return rid.value

    def next_frame(self, timeout_ms=1000, want_rg8=True, copy=True):
        """(info, rg8 array or None), or None on timeout.  copy=False returns a
        view of the server object's frame buffer, valid until the next call."""
        info = FrameInfo()
        while True:
            if want_rg8:
                buf = self._buf
                ptr = buf.ctypes.data_as(C.POINTER(C.c_uint8))
                n = buf.size
            else:
                buf, ptr, n = None, None, 0
            rc = self.L.lbm_server_next_frame(self.h, int(timeout_ms), C.byref(info), ptr, n)
            if rc == -5 and want_rg8 and info.bytes > n:  # LBM_ERR_CAPACITY: grow, retry
                self._buf = np.empty(int(info.bytes), np.uint8)
                continue
            self._ck(rc, "lbm_server_next_frame")
            break
        if rc == 0:
            return None
        if not want_rg8:
            return info, None
        return info, (buf[:info.bytes].copy() if copy else buf[:info.bytes])

    def stats(self):
        st = ServerStats()
        self._ck(self.L.lbm_server_stats_get(self.h, C.byref(st)), "lbm_server_stats_get")
        return st.as_dict()

    def clock(self):
        return self.L.lbm_server_clock(self.h)

    def pause(self, paused=True):
        self._ck(self.L.lbm_server_pause(self.h, 1 if paused else 0), "lbm_server_pause")


def state_bytes(nx, ny, nz, storage="f32"):
    return lib().lbm_state_bytes(nx, ny, nz, STORAGE[storage])

