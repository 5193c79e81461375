// N-API addon over liblbm's C ABI (include/lbm.h).  A thin argument marshaller:
// typed arrays are passed by pointer, integers and doubles by value, and any
// nonzero status becomes a thrown JS Error carrying lbm_last_error().  The JS
// modules sim.js / solid.js in this directory build the reference's Sim and
// RigidSphere surface (src/sim.js, src/solid.js) on top of it.
#include <node_api.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "lbm.h"

namespace {

#define NAPI_OK(env, call)                                             \
  do {                                                                 \
    if ((call) != napi_ok) {                                           \
      napi_throw_error((env), nullptr, "N-API call failed: " #call);   \
      return nullptr;                                                  \
    }                                                                  \
  } while (0)

struct Args {
  napi_value v[10];
  size_t n = 10;
  napi_value self;
};

bool get_args(napi_env env, napi_callback_info info, Args &a) {
  a.n = 10;
  return napi_get_cb_info(env, info, &a.n, a.v, &a.self, nullptr) == napi_ok;
}

napi_value throw_status(napi_env env, lbm_sim *s, int rc, const char *what) {
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + lbm_last_error(s);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

// The external's data: the handle, cleared by destroy() so the finalizer and
// later calls see it is gone.
struct Holder {
  lbm_sim *s;
};

lbm_sim *get_sim(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a liblbm handle");
    return nullptr;
  }
  lbm_sim *s = static_cast<Holder *>(p)->s;
  if (!s) napi_throw_error(env, nullptr, "the simulation was disposed");
  return s;
}

// Typed-array data pointer and element count; nullptr for null/undefined.
void *typed_data(napi_env env, napi_value v, size_t *len, napi_typedarray_type *type) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_null || t == napi_undefined) return nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return nullptr;
  void *data = nullptr;
  napi_value ab;
  size_t off = 0;
  napi_get_typedarray_info(env, v, type, len, &data, &ab, &off);
  return data;
}

lbm_sphere *get_sphere(napi_env env, napi_value v) {
  size_t len = 0;
  napi_typedarray_type type;
  void *p = typed_data(env, v, &len, &type);
  if (!p) return nullptr;
  if (type != napi_float64_array || len < sizeof(lbm_sphere) / sizeof(double)) {
    napi_throw_type_error(env, nullptr, "sphere state must be a Float64Array(27)");
    return nullptr;
  }
  return static_cast<lbm_sphere *>(p);
}

double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

int32_t i32(napi_env env, napi_value v) {
  int32_t d = 0;
  napi_get_value_int32(env, v, &d);
  return d;
}

napi_value undef(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

void finalize_sim(napi_env, void *data, void *) {
  Holder *h = static_cast<Holder *>(data);
  if (h->s) lbm_destroy(h->s);
  delete h;
}

// create(nx, ny, nz) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 3) {
    napi_throw_type_error(env, nullptr, "create(nx, ny, nz)");
    return nullptr;
  }
  lbm_sim *s = nullptr;
  const int rc = lbm_create(i32(env, a.v[0]), i32(env, a.v[1]), i32(env, a.v[2]), &s);
  if (rc) return throw_status(env, nullptr, rc, "lbm_create");
  napi_value h;
  Holder *holder = new Holder{s};
  if (napi_create_external(env, holder, finalize_sim, nullptr, &h) != napi_ok) {
    lbm_destroy(s);
    delete holder;
    napi_throw_error(env, nullptr, "napi_create_external failed");
    return nullptr;
  }
  return h;
}

// destroy(handle): frees the device state now (the finalizer then has nothing to do).
napi_value Destroy(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  void *p = nullptr;
  if (napi_get_value_external(env, a.v[0], &p) == napi_ok && p) {
    Holder *h = static_cast<Holder *>(p);
    if (h->s) lbm_destroy(h->s);
    h->s = nullptr;
  }
  return undef(env);
}

// setParams(h, Float64Array[8]) in lbm_params order.
napi_value SetParams(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  size_t len = 0;
  napi_typedarray_type t;
  double *p = static_cast<double *>(typed_data(env, a.v[1], &len, &t));
  if (!p || t != napi_float64_array || len < 8) {
    napi_throw_type_error(env, nullptr, "setParams(h, Float64Array(8))");
    return nullptr;
  }
  lbm_params q{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
  const int rc = lbm_set_params(s, &q);
  if (rc) return throw_status(env, s, rc, "lbm_set_params");
  return undef(env);
}

#define SIM_NOARG(Name, fn)                                   \
  napi_value Name(napi_env env, napi_callback_info info) {    \
    Args a;                                                   \
    get_args(env, info, a);                                   \
    lbm_sim *s = get_sim(env, a.v[0]);                        \
    if (!s) return nullptr;                                   \
    const int rc = fn(s);                                     \
    if (rc) return throw_status(env, s, rc, #fn);             \
    return undef(env);                                        \
  }

SIM_NOARG(SeedHydrostatic, lbm_seed_hydrostatic)
SIM_NOARG(HandleFresh, lbm_handle_fresh_cells)
SIM_NOARG(Gradients, lbm_compute_phi_gradients)
SIM_NOARG(CollidePhase, lbm_collide_phase)
SIM_NOARG(StreamPhase, lbm_stream_phase)
SIM_NOARG(CollideHydro, lbm_collide_hydro)
SIM_NOARG(Synchronize, lbm_synchronize)

#define SIM_SPHERE(Name, fn)                                  \
  napi_value Name(napi_env env, napi_callback_info info) {    \
    Args a;                                                   \
    get_args(env, info, a);                                   \
    lbm_sim *s = get_sim(env, a.v[0]);                        \
    if (!s) return nullptr;                                   \
    lbm_sphere *b = a.n > 1 ? get_sphere(env, a.v[1]) : nullptr; \
    const int rc = fn(s, b);                                  \
    if (rc) return throw_status(env, s, rc, #fn);             \
    return undef(env);                                        \
  }

SIM_SPHERE(Retag, lbm_retag_cells)
SIM_SPHERE(StreamHydroMacro, lbm_stream_hydro_macro)
SIM_SPHERE(SimStep, lbm_sim_step)
SIM_SPHERE(AbsorbDead, lbm_absorb_dead_cells)

// initFlat(h, y)
napi_value InitFlat(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  const int rc = lbm_init_flat(s, num(env, a.v[1]));
  if (rc) return throw_status(env, s, rc, "lbm_init_flat");
  return undef(env);
}

// initWaterlineMap(h, Float64Array(nx*nz))
napi_value InitWaterlineMap(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  size_t len = 0;
  napi_typedarray_type t;
  double *p = static_cast<double *>(typed_data(env, a.v[1], &len, &t));
  if (!p || t != napi_float64_array) {
    napi_throw_type_error(env, nullptr, "initWaterlineMap(h, Float64Array)");
    return nullptr;
  }
  const int rc = lbm_init_waterline_map(s, p);
  if (rc) return throw_status(env, s, rc, "lbm_init_waterline_map");
  return undef(env);
}

// retagMask(h, Uint8Array(n), Float32Array(3n))
napi_value RetagMask(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  size_t l1 = 0, l2 = 0;
  napi_typedarray_type t1, t2;
  uint8_t *m = static_cast<uint8_t *>(typed_data(env, a.v[1], &l1, &t1));
  float *us = static_cast<float *>(typed_data(env, a.v[2], &l2, &t2));
  if (!m || !us || t1 != napi_uint8_array || t2 != napi_float32_array) {
    napi_throw_type_error(env, nullptr, "retagMask(h, Uint8Array, Float32Array)");
    return nullptr;
  }
  const int rc = lbm_retag_cells_mask(s, m, us);
  if (rc) return throw_status(env, s, rc, "lbm_retag_cells_mask");
  return undef(env);
}

// streamHydroMacroLinks(h) -> Float64Array of (x, y, z, k, px, py, pz) records
napi_value StreamHydroMacroLinks(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  int count = 0;
  int rc = lbm_stream_hydro_macro_links(s, &count);
  if (rc) return throw_status(env, s, rc, "lbm_stream_hydro_macro_links");
  std::vector<lbm_link> links(count > 0 ? count : 1);
  rc = lbm_get_links(s, links.data(), count);
  if (rc) return throw_status(env, s, rc, "lbm_get_links");
  napi_value ab, out;
  void *data = nullptr;
  NAPI_OK(env, napi_create_arraybuffer(env, (size_t)count * 7 * sizeof(double), &data, &ab));
  double *d = static_cast<double *>(data);
  for (int j = 0; j < count; j++) {
    d[7 * j + 0] = links[j].x;
    d[7 * j + 1] = links[j].y;
    d[7 * j + 2] = links[j].z;
    d[7 * j + 3] = links[j].k;
    d[7 * j + 4] = links[j].px;
    d[7 * j + 5] = links[j].py;
    d[7 * j + 6] = links[j].pz;
  }
  NAPI_OK(env, napi_create_typedarray(env, napi_float64_array, (size_t)count * 7, ab, 0, &out));
  return out;
}

// step(h, sphere | null, nsteps): fused coupled steps.
napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  lbm_sphere *b = get_sphere(env, a.v[1]);
  const int rc = lbm_step(s, b, i32(env, a.v[2]));
  if (rc) return throw_status(env, s, rc, "lbm_step");
  return undef(env);
}

// stepAsync(h, sphere | null, nsteps) -> Promise: lbm_step on a worker thread,
// so the event loop keeps running while the GPU steps (a serving process).
// The handle and the sphere array stay referenced until the step completes;
// the caller must not use the simulation meanwhile.
struct StepWork {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref handle_ref = nullptr, sphere_ref = nullptr;
  lbm_sim *s = nullptr;
  lbm_sphere *b = nullptr;
  int nsteps = 0;
  int rc = 0;
  std::string err;
};

void step_execute(napi_env, void *data) {
  StepWork *w = static_cast<StepWork *>(data);
  w->rc = lbm_step(w->s, w->b, w->nsteps);
  if (w->rc) w->err = lbm_last_error(w->s);
}

void step_complete(napi_env env, napi_status, void *data) {
  StepWork *w = static_cast<StepWork *>(data);
  if (w->rc == 0) {
    napi_resolve_deferred(env, w->deferred, undef(env));
  } else {
    const std::string msg = "lbm_step failed (" + std::to_string(w->rc) + "): " + w->err;
    napi_value m, e;
    napi_create_string_utf8(env, msg.c_str(), msg.size(), &m);
    napi_create_error(env, nullptr, m, &e);
    napi_reject_deferred(env, w->deferred, e);
  }
  napi_delete_reference(env, w->handle_ref);
  if (w->sphere_ref) napi_delete_reference(env, w->sphere_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value StepAsync(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  StepWork *w = new StepWork();
  w->s = s;
  w->b = get_sphere(env, a.v[1]);
  w->nsteps = i32(env, a.v[2]);
  napi_value promise, name;
  NAPI_OK(env, napi_create_promise(env, &w->deferred, &promise));
  NAPI_OK(env, napi_create_reference(env, a.v[0], 1, &w->handle_ref));
  if (w->b) NAPI_OK(env, napi_create_reference(env, a.v[1], 1, &w->sphere_ref));
  NAPI_OK(env, napi_create_string_utf8(env, "lbm_step", NAPI_AUTO_LENGTH, &name));
  NAPI_OK(env, napi_create_async_work(env, nullptr, name, step_execute, step_complete, w, &w->work));
  NAPI_OK(env, napi_queue_async_work(env, w->work));
  return promise;
}

// checkFields(h) -> [allFinite, maxU]
napi_value CheckFields(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  int ok = 0;
  double um = 0;
  const int rc = lbm_check_fields(s, &ok, &um);
  if (rc) return throw_status(env, s, rc, "lbm_check_fields");
  napi_value arr, v0, v1;
  NAPI_OK(env, napi_create_array_with_length(env, 2, &arr));
  NAPI_OK(env, napi_get_boolean(env, ok != 0, &v0));
  NAPI_OK(env, napi_create_double(env, um, &v1));
  NAPI_OK(env, napi_set_element(env, arr, 0, v0));
  NAPI_OK(env, napi_set_element(env, arr, 1, v1));
  return arr;
}

// download(h, field, typedArray) / upload(h, field, typedArray)
napi_value Transfer(napi_env env, napi_callback_info info, bool down) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  const int field = i32(env, a.v[1]);
  size_t len = 0;
  napi_typedarray_type t;
  void *p = typed_data(env, a.v[2], &len, &t);
  const int64_t want = lbm_field_count(s, field);
  if (!p || want < 0 || (int64_t)len != want) {
    napi_throw_range_error(env, nullptr, "field transfer: typed array of the field's length required");
    return nullptr;
  }
  const int rc = down ? lbm_download(s, field, p) : lbm_upload(s, field, p);
  if (rc) return throw_status(env, s, rc, down ? "lbm_download" : "lbm_upload");
  return undef(env);
}
napi_value Download(napi_env env, napi_callback_info info) { return Transfer(env, info, true); }
napi_value Upload(napi_env env, napi_callback_info info) { return Transfer(env, info, false); }

// pack(h, Uint8Array(2n))
napi_value Pack(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  size_t len = 0;
  napi_typedarray_type t;
  uint8_t *p = static_cast<uint8_t *>(typed_data(env, a.v[1], &len, &t));
  if (!p || (int64_t)len < 2 * lbm_field_count(s, LBM_FIELD_PHI)) {
    napi_throw_range_error(env, nullptr, "packVolume needs a Uint8Array of 2n bytes");
    return nullptr;
  }
  const int rc = lbm_pack_volume(s, p);
  if (rc) return throw_status(env, s, rc, "lbm_pack_volume");
  return undef(env);
}

napi_value CellList(napi_env env, napi_callback_info info, bool fresh) {
  Args a;
  get_args(env, info, a);
  lbm_sim *s = get_sim(env, a.v[0]);
  if (!s) return nullptr;
  int count = 0;
  int rc = fresh ? lbm_fresh_cells(s, nullptr, 0, &count) : lbm_dead_cells(s, nullptr, 0, &count);
  if (rc) return throw_status(env, s, rc, "cell list");
  napi_value ab, out;
  void *data = nullptr;
  NAPI_OK(env, napi_create_arraybuffer(env, (size_t)count * 4, &data, &ab));
  if (count) {
    rc = fresh ? lbm_fresh_cells(s, static_cast<int32_t *>(data), count, &count)
               : lbm_dead_cells(s, static_cast<int32_t *>(data), count, &count);
    if (rc) return throw_status(env, s, rc, "cell list");
  }
  NAPI_OK(env, napi_create_typedarray(env, napi_int32_array, (size_t)count, ab, 0, &out));
  return out;
}
napi_value FreshCells(napi_env env, napi_callback_info info) { return CellList(env, info, true); }
napi_value DeadCells(napi_env env, napi_callback_info info) { return CellList(env, info, false); }

// ---------------------------------------------------------------- sphere

// sphereInit(Float64Array(27), nx, ny, nz, cx, cy, cz, r, density): NaN = default
napi_value SphereInit(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  lbm_sphere_init(b, i32(env, a.v[1]), i32(env, a.v[2]), i32(env, a.v[3]), num(env, a.v[4]),
                  num(env, a.v[5]), num(env, a.v[6]), num(env, a.v[7]), num(env, a.v[8]));
  return undef(env);
}

napi_value SphereRecomputeMass(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  lbm_sphere_recompute_mass(b);
  return undef(env);
}

napi_value SphereContains(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  napi_value r;
  NAPI_OK(env, napi_get_boolean(env, lbm_sphere_contains(b, i32(env, a.v[1]), i32(env, a.v[2]),
                                                         i32(env, a.v[3])) != 0, &r));
  return r;
}

// sphereSurfaceVelocity(state, x, y, z, out): writes out[0..2], returns out
napi_value SphereSurfaceVelocity(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  double v[3];
  lbm_sphere_surface_velocity(b, i32(env, a.v[1]), i32(env, a.v[2]), i32(env, a.v[3]), v);
  for (uint32_t k = 0; k < 3; k++) {
    napi_value d;
    NAPI_OK(env, napi_create_double(env, v[k], &d));
    NAPI_OK(env, napi_set_element(env, a.v[4], k, d));
  }
  return a.v[4];
}

napi_value SphereApplyImpulse(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  lbm_sphere_apply_impulse_at_cell(b, i32(env, a.v[1]), i32(env, a.v[2]), i32(env, a.v[3]),
                                   num(env, a.v[4]), num(env, a.v[5]), num(env, a.v[6]));
  return undef(env);
}

napi_value SphereStep(napi_env env, napi_callback_info info) {
  Args a;
  get_args(env, info, a);
  lbm_sphere *b = get_sphere(env, a.v[0]);
  if (!b) return nullptr;
  lbm_sphere_step(b, num(env, a.v[1]), num(env, a.v[2]));
  return undef(env);
}

// latticeTables() -> {EX27, EY27, EZ27, W27, OPP27, EX7, EY7, EZ7, W7, OPP7}
napi_value LatticeTables(napi_env env, napi_callback_info) {
  int32_t ex[27], ey[27], ez[27], opp[27], ex7[7], ey7[7], ez7[7], opp7[7];
  float w[27], w7[7];
  lbm_lattice_tables(ex, ey, ez, w, opp, ex7, ey7, ez7, w7, opp7);
  napi_value obj;
  NAPI_OK(env, napi_create_object(env, &obj));
  auto put_i = [&](const char *name, const int32_t *v, int n) {
    napi_value arr;
    napi_create_array_with_length(env, n, &arr);
    for (int k = 0; k < n; k++) {
      napi_value e;
      napi_create_int32(env, v[k], &e);
      napi_set_element(env, arr, k, e);
    }
    napi_set_named_property(env, obj, name, arr);
  };
  auto put_f = [&](const char *name, const float *v, int n) {
    napi_value arr;
    napi_create_array_with_length(env, n, &arr);
    for (int k = 0; k < n; k++) {
      napi_value e;
      napi_create_double(env, (double)v[k], &e);
      napi_set_element(env, arr, k, e);
    }
    napi_set_named_property(env, obj, name, arr);
  };
  put_i("EX27", ex, 27);
  put_i("EY27", ey, 27);
  put_i("EZ27", ez, 27);
  put_i("OPP27", opp, 27);
  put_f("W27", w, 27);
  put_i("EX7", ex7, 7);
  put_i("EY7", ey7, 7);
  put_i("EZ7", ez7, 7);
  put_i("OPP7", opp7, 7);
  put_f("W7", w7, 7);
  return obj;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(env, napi_create_string_utf8(env, lbm_version(), NAPI_AUTO_LENGTH, &v));
  return v;
}

}  // namespace

NAPI_MODULE_INIT() {
  const napi_property_descriptor props[] = {
#define FN(name, cb) {name, nullptr, cb, nullptr, nullptr, nullptr, napi_default, nullptr}
      FN("create", Create),
      FN("destroy", Destroy),
      FN("setParams", SetParams),
      FN("initFlat", InitFlat),
      FN("initWaterlineMap", InitWaterlineMap),
      FN("seedHydrostatic", SeedHydrostatic),
      FN("retag", Retag),
      FN("retagMask", RetagMask),
      FN("handleFresh", HandleFresh),
      FN("gradients", Gradients),
      FN("collidePhase", CollidePhase),
      FN("streamPhase", StreamPhase),
      FN("collideHydro", CollideHydro),
      FN("streamHydroMacro", StreamHydroMacro),
      FN("streamHydroMacroLinks", StreamHydroMacroLinks),
      FN("simStep", SimStep),
      FN("absorbDead", AbsorbDead),
      FN("step", Step),
      FN("stepAsync", StepAsync),
      FN("checkFields", CheckFields),
      FN("download", Download),
      FN("upload", Upload),
      FN("pack", Pack),
      FN("freshCells", FreshCells),
      FN("deadCells", DeadCells),
      FN("synchronize", Synchronize),
      FN("sphereInit", SphereInit),
      FN("sphereRecomputeMass", SphereRecomputeMass),
      FN("sphereContains", SphereContains),
      FN("sphereSurfaceVelocity", SphereSurfaceVelocity),
      FN("sphereApplyImpulse", SphereApplyImpulse),
      FN("sphereStep", SphereStep),
      FN("latticeTables", LatticeTables),
      FN("version", Version),
#undef FN
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}
