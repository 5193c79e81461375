"""ctypes binding of the CPU oracle (oracle/_build/liboracle{32,64}.so).

Test infrastructure only: the oracle is the checker for the HIP library."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
BUILD = os.path.join(ORACLE_DIR, "_build")

SPHERE_FIELDS = ["nx", "ny", "nz", "cx", "cy", "cz", "r", "vx", "vy", "vz", "wx", "wy", "wz",
                 "density", "volume", "mass", "I", "dragging", "dragTargetX", "dragTargetY",
                 "dragTargetZ", "Fx", "Fy", "Fz", "Tx", "Ty", "Tz"]


class OSphere(C.Structure):
    """Mirror of RigidSphere's fields (all doubles)."""
    _fields_ = [(k, C.c_double) for k in SPHERE_FIELDS]

    def as_list(self):
        return [getattr(self, k) for k in SPHERE_FIELDS]

    def load_list(self, vals):
        for k, v in zip(SPHERE_FIELDS, vals):
            setattr(self, k, float(v))


def _osim_struct(store):
    P = C.POINTER(store)
    U8 = C.POINTER(C.c_uint8)
    fields = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n", C.c_int)]
    fields += [(k, P) for k in ("f", "f2", "h", "h2", "rho", "ux", "uy", "uz", "phi", "phiPrev")]
    fields += [("tag", U8), ("tagPrev", U8)]
    fields += [(k, P) for k in ("usx", "usy", "usz", "gx", "gy", "gz", "lap",
                                 "bodyFx", "bodyFy", "bodyFz")]
    fields += [("neighOffset27", C.c_int32 * 27), ("neighOffset7", C.c_int32 * 7)]
    fields += [(k, C.c_double) for k in ("tau", "tauPhi", "sigma", "gravity", "rhoL", "rhoG",
                                          "W", "mobility")]
    return type("OSimStruct", (C.Structure,), {"_fields_": fields})


_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(bits=32):
    """bits: 32 (float32 storage, the reference's), 64 (FP64 storage), or
    "32_omp" / "64_omp" (those builds with their per-cell passes in OpenMP
    threads, same results)."""
    if bits in _LIBS:
        return _LIBS[bits]
    path = os.path.join(BUILD, f"liboracle{bits}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    fp64 = bits in (64, "64_omp")
    store = C.c_double if fp64 else C.c_float
    S = _osim_struct(store)
    SP = C.POINTER(S)
    B = C.POINTER(OSphere)
    L.oracle_sim_create.restype = SP
    L.oracle_sim_create.argtypes = [C.c_int, C.c_int, C.c_int]
    L.oracle_sim_destroy.argtypes = [SP]
    L.oracle_init_flat.argtypes = [SP, C.c_double]
    L.oracle_init_waterline_map.argtypes = [SP, C.POINTER(C.c_double)]
    for fn in ("oracle_seed_hydrostatic", "oracle_compute_phi_gradients",
               "oracle_handle_fresh_cells", "oracle_collide_phase", "oracle_stream_phase",
               "oracle_collide_hydro"):
        getattr(L, fn).argtypes = [SP]
    L.oracle_retag_cells.argtypes = [SP, B]
    L.oracle_stream_hydro_and_macro.argtypes = [SP, B]
    L.oracle_sim_step.argtypes = [SP, B]
    L.oracle_pack_volume.argtypes = [SP, C.POINTER(C.c_uint8)]
    L.oracle_sphere_init.argtypes = [B, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double]
    L.oracle_sphere_recompute_mass.argtypes = [B]
    L.oracle_sphere_contains.argtypes = [B, C.c_int, C.c_int, C.c_int]
    L.oracle_sphere_contains.restype = C.c_int
    L.oracle_sphere_surface_velocity.argtypes = [B, C.c_int, C.c_int, C.c_int,
                                                 C.POINTER(C.c_double)]
    L.oracle_sphere_apply_impulse_at_cell.argtypes = [B, C.c_int, C.c_int, C.c_int, C.c_double,
                                                      C.c_double, C.c_double]
    L.oracle_sphere_absorb_dead_cells.argtypes = [B, SP]
    L.oracle_sphere_step.argtypes = [B, C.c_double, C.c_double]
    L.oracle_coupled_steps.argtypes = [SP, B, C.c_int]
    L.oracle_storage_bytes.restype = C.c_int
    _LIBS[bits] = (L, S, store)
    return _LIBS[bits]


POP = {"f": 27, "f2": 27, "h": 7, "h2": 7}
FIELDS = ["f", "f2", "h", "h2", "rho", "ux", "uy", "uz", "phi", "phiPrev", "usx", "usy", "usz",
          "gx", "gy", "gz", "lap", "bodyFx", "bodyFy", "bodyFz"]
PARAMS = ["tau", "tauPhi", "sigma", "gravity", "rhoL", "rhoG", "W", "mobility"]


class OSim:
    """Oracle Sim.  Arrays are numpy views on the oracle's memory, in the
    reference's layout (f[i*27+k])."""

    def __init__(self, nx, ny, nz, bits=32):
        self.L, self.S, store = lib(bits)
        self.ptr = self.L.oracle_sim_create(nx, ny, nz)
        self.s = self.ptr.contents
        self.nx, self.ny, self.nz, self.n = nx, ny, nz, nx * ny * nz
        dt = np.float64 if bits in (64, "64_omp") else np.float32
        for k in FIELDS:
            cnt = self.n * POP.get(k, 1)
            arr = np.ctypeslib.as_array(getattr(self.s, k), shape=(cnt,))
            assert arr.dtype == dt
            setattr(self, k, arr)
        self.tag = np.ctypeslib.as_array(self.s.tag, shape=(self.n,))
        self.tagPrev = np.ctypeslib.as_array(self.s.tagPrev, shape=(self.n,))

    def __del__(self):
        try:
            self.L.oracle_sim_destroy(self.ptr)
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            assert k in PARAMS, k
            setattr(self.s, k, float(v))

    def params(self):
        return {k: getattr(self.s, k) for k in PARAMS}

    def init_flat(self, y):
        self.L.oracle_init_flat(self.ptr, float(y))

    def init_waterline_map(self, yw):
        yw = np.ascontiguousarray(yw, dtype=np.float64)
        assert yw.size == self.nx * self.nz
        self.L.oracle_init_waterline_map(self.ptr, yw.ctypes.data_as(C.POINTER(C.c_double)))

    def seed_hydrostatic(self):
        self.L.oracle_seed_hydrostatic(self.ptr)

    def retag(self, sphere):
        self.L.oracle_retag_cells(self.ptr, C.byref(sphere) if sphere is not None else None)

    def fresh(self):
        self.L.oracle_handle_fresh_cells(self.ptr)

    def gradients(self):
        self.L.oracle_compute_phi_gradients(self.ptr)

    def collide_phase(self):
        self.L.oracle_collide_phase(self.ptr)

    def stream_phase(self):
        self.L.oracle_stream_phase(self.ptr)

    def collide_hydro(self):
        self.L.oracle_collide_hydro(self.ptr)

    def stream_hydro_macro(self, sphere):
        self.L.oracle_stream_hydro_and_macro(self.ptr, C.byref(sphere) if sphere is not None else None)

    def step(self, sphere):
        self.L.oracle_sim_step(self.ptr, C.byref(sphere) if sphere is not None else None)

    def coupled(self, sphere, nsteps):
        self.L.oracle_coupled_steps(self.ptr, C.byref(sphere) if sphere is not None else None, nsteps)

    def pack(self):
        out = np.zeros(2 * self.n, np.uint8)
        self.L.oracle_pack_volume(self.ptr, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    def absorb_dead(self, sphere):
        self.L.oracle_sphere_absorb_dead_cells(C.byref(sphere), self.ptr)


def sphere(nx, ny, nz, cx, cy, cz, r, density, bits=32):
    L = lib(bits)[0]
    b = OSphere()
    L.oracle_sphere_init(C.byref(b), nx, ny, nz, cx, cy, cz, r, density)
    return b


def sphere_step(b, dt, g, bits=32):
    lib(bits)[0].oracle_sphere_step(C.byref(b), dt, g)
