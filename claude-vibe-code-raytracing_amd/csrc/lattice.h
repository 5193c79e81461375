// Lattice tables and the rigid-sphere functions, shared by host and device code.
//
// Tables restate src/lattices.js: D3Q27 in opposite pairs (k odd <-> k+1), its
// weights rounded to float32 as the reference's Float32Array stores them, and
// D3Q7 with the same pairing.  Loops over k are fully unrolled in the kernels,
// so every table entry becomes an immediate.
//
// The sphere functions restate src/solid.js.  They are __host__ __device__ so
// the host path (per-pass API) and the device integrator (fused coupled step)
// run identical FP64 code; the library is built with -ffp-contract=off on both
// sides, so the two agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "lbm.h"

namespace lbm {

constexpr int Q27 = 27;
constexpr int Q7 = 7;
constexpr uint8_t TAG_FLUID = LBM_TAG_FLUID;
constexpr uint8_t TAG_SOLID = LBM_TAG_SOLID;
constexpr uint8_t TAG_WALL = LBM_TAG_WALL;

// src/lattices.js:34-49
constexpr int EX27[Q27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                           1, -1, 1, -1, 1, -1, -1, 1};
constexpr int EY27[Q27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                           1, -1, 1, -1, -1, 1, 1, -1};
constexpr int EZ27[Q27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                           1, -1, -1, 1, 1, -1, 1, -1};
// The components of e_k packed two bits each (component + 1), so that a
// direction chosen at run time is decoded without a table in memory.
constexpr uint64_t pack_dirs(const int (&e)[Q27]) {
  uint64_t v = 0;
  for (int k = 0; k < Q27; k++) v |= (uint64_t)(e[k] + 1) << (2 * k);
  return v;
}
constexpr uint64_t kEX27Bits = pack_dirs(EX27), kEY27Bits = pack_dirs(EY27),
                   kEZ27Bits = pack_dirs(EZ27);
__host__ __device__ inline int dir_comp(uint64_t bits, int k) {
  return (int)((bits >> (2 * k)) & 3u) - 1;
}
// src/lattices.js:66-69
__host__ __device__ constexpr int opp27(int k) { return k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1); }
// src/lattices.js:52-57, stored as float32 (lattices.js:22).
__host__ __device__ constexpr float w27(int k) {
  return k == 0 ? (float)(8.0 / 27) : k <= 6 ? (float)(2.0 / 27) : k <= 18 ? (float)(1.0 / 54)
                                                                            : (float)(1.0 / 216);
}
// src/lattices.js:79-91
constexpr int EX7[Q7] = {0, 1, -1, 0, 0, 0, 0};
constexpr int EY7[Q7] = {0, 0, 0, 1, -1, 0, 0};
constexpr int EZ7[Q7] = {0, 0, 0, 0, 0, 1, -1};
__host__ __device__ constexpr int opp7(int k) { return k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1); }
__host__ __device__ constexpr float w7(int k) { return k == 0 ? 0.25f : 0.125f; }

// src/lattices.js:95-99
constexpr double INV_CS2_27 = 3.0;
constexpr double INV_2CS4_27 = 4.5;
constexpr double CS2_7 = 1.0 / 4;
constexpr double INV_CS2_7 = 4.0;

// ---------------------------------------------------------------- sphere

// RigidSphere.contains (src/solid.js:50-55).
__host__ __device__ inline bool sphere_contains(const lbm_sphere &b, int x, int y, int z) {
  const double dx = x + 0.5 - b.cx;
  const double dy = y + 0.5 - b.cy;
  const double dz = z + 0.5 - b.cz;
  return dx * dx + dy * dy + dz * dz <= b.r * b.r;
}

// RigidSphere.surfaceVelocity (src/solid.js:57-66).
__host__ __device__ inline void sphere_surface_velocity(const lbm_sphere &b, int x, int y, int z,
                                                        double out[3]) {
  const double rx = x + 0.5 - b.cx;
  const double ry = y + 0.5 - b.cy;
  const double rz = z + 0.5 - b.cz;
  out[0] = b.vx + (b.wy * rz - b.wz * ry);
  out[1] = b.vy + (b.wz * rx - b.wx * rz);
  out[2] = b.vz + (b.wx * ry - b.wy * rx);
}

// The six increments RigidSphere.applyImpulseAtCell (src/solid.js:70-80) adds
// to Fx, Fy, Fz, Tx, Ty, Tz: the impulse p and the torque r x p about the centre.
// They depend on the centre only, so a run of impulses may have its increments
// computed in any order and summed one after the other.
__host__ __device__ inline void impulse_terms(const lbm_sphere &b, int x, int y, int z, double px,
                                              double py, double pz, double t[6]) {
  const double rx = x + 0.5 - b.cx;
  const double ry = y + 0.5 - b.cy;
  const double rz = z + 0.5 - b.cz;
  t[0] = px;
  t[1] = py;
  t[2] = pz;
  t[3] = ry * pz - rz * py;
  t[4] = rz * px - rx * pz;
  t[5] = rx * py - ry * px;
}

// RigidSphere.applyImpulseAtCell (src/solid.js:70-80).
__host__ __device__ inline void sphere_apply_impulse(lbm_sphere &b, int x, int y, int z, double px,
                                                     double py, double pz) {
  double t[6];
  impulse_terms(b, x, y, z, px, py, pz, t);
  b.Fx += t[0];
  b.Fy += t[1];
  b.Fz += t[2];
  b.Tx += t[3];
  b.Ty += t[4];
  b.Tz += t[5];
}

// RigidSphere.recomputeMass (src/solid.js:39-44).
__host__ __device__ inline void sphere_recompute_mass(lbm_sphere &b) {
  const double PI = 3.141592653589793;  // Math.PI
  const double V = (4.0 / 3) * PI * b.r * b.r * b.r;
  b.volume = V;
  b.mass = b.density * V;
  b.I = (2.0 / 5) * b.mass * b.r * b.r;
}

// Math.max(0, v) and Math.min(0, v), NaN and signed-zero rules included.
__host__ __device__ inline double js_max0(double v) { return v != v ? v : (v > 0 ? v : 0.0); }
__host__ __device__ inline double js_min0(double v) {
  if (v != v) return v;
  if (v < 0) return v;
  return (v == 0 && signbit(v)) ? -0.0 : 0.0;
}

// RigidSphere.step(dt, gravity) (src/solid.js:104-210), in the reference's order:
// gravity, drag spring, capped linear acceleration, angular update, speed cap,
// damping, snap to rest, move, clamp inside the box, reset accumulators.
__host__ __device__ inline void sphere_step(lbm_sphere &b, double dt, double gravity) {
  b.Fy += b.mass * gravity;
  if (b.dragging != 0) {
    const double k = 0.004 * b.mass;
    const double cD = 2.0 * sqrt(k * b.mass);
    b.Fx += k * (b.dragTargetX - b.cx) - cD * b.vx;
    b.Fy += k * (b.dragTargetY - b.cy) - cD * b.vy;
    b.Fz += k * (b.dragTargetZ - b.cz) - cD * b.vz;
  }
  // Math.abs(gravity || 0.001): 0 and NaN select 0.001.
  const double gsel = (gravity != 0 && gravity == gravity) ? gravity : 0.001;
  const double accCap = 3.5 * fabs(gsel);
  {
    const double ax = b.Fx / b.mass;
    const double ay = b.Fy / b.mass;
    const double az = b.Fz / b.mass;
    const double a2 = ax * ax + ay * ay + az * az;
    const double c2 = accCap * accCap;
    const double s = a2 > c2 ? accCap / sqrt(a2) : 1.0;
    b.vx += dt * ax * s;
    b.vy += dt * ay * s;
    b.vz += dt * az * s;
  }
  b.wx += dt * b.Tx / b.I;
  b.wy += dt * b.Ty / b.I;
  b.wz += dt * b.Tz / b.I;
  const double BODY_UMAX = 0.08;
  const double sp2 = b.vx * b.vx + b.vy * b.vy + b.vz * b.vz;
  if (sp2 > BODY_UMAX * BODY_UMAX) {
    const double s = BODY_UMAX / sqrt(sp2);
    b.vx *= s;
    b.vy *= s;
    b.vz *= s;
  }
  const double linDamp = 0.98;
  const double omegaDamp = 0.95;
  b.vx *= linDamp;
  b.vy *= linDamp;
  b.vz *= linDamp;
  b.wx *= omegaDamp;
  b.wy *= omegaDamp;
  b.wz *= omegaDamp;
  const double SNAP = 0.0001;
  if (fabs(b.vx) < SNAP) b.vx = 0;
  if (fabs(b.vy) < SNAP) b.vy = 0;
  if (fabs(b.vz) < SNAP) b.vz = 0;
  if (fabs(b.wx) < SNAP) b.wx = 0;
  if (fabs(b.wy) < SNAP) b.wy = 0;
  if (fabs(b.wz) < SNAP) b.wz = 0;
  b.cx += dt * b.vx;
  b.cy += dt * b.vy;
  b.cz += dt * b.vz;
  const double m = b.r + 1.5;
  if (b.cx < m) { b.cx = m; b.vx = js_max0(b.vx); }
  if (b.cx > b.nx - 1 - m) { b.cx = b.nx - 1 - m; b.vx = js_min0(b.vx); }
  if (b.cy < m) { b.cy = m; b.vy = js_max0(b.vy); }
  if (b.cy > b.ny - 1 - m) { b.cy = b.ny - 1 - m; b.vy = js_min0(b.vy); }
  if (b.cz < m) { b.cz = m; b.vz = js_max0(b.vz); }
  if (b.cz > b.nz - 1 - m) { b.cz = b.nz - 1 - m; b.vz = js_min0(b.vz); }
  b.Fx = b.Fy = b.Fz = 0;
  b.Tx = b.Ty = b.Tz = 0;
}

}  // namespace lbm
