// Kernels of the fused coupled step (lbm_step).
//
// The reference's step t runs retag(t), fresh(t), gradients, collide(t),
// stream(t) and the macroscopic update, then the body absorbs dead cells and
// moves.  The fused form reorders nothing; it regroups the passes around the
// only global dependency, the sphere position, which needs the momentum sum of
// stream(t) before retag(t+1) can run:
//
//   km_phase_stream(t)    phi(t+1) = clamp(sum of pulled h) for every cell
//   km_stream_collide(t)  stream(t) + macro(t) for every fluid cell, then
//                         collide(t+1) for every fluid cell outside the band;
//                         band cells keep their post-stream populations
//   km_body(t)            momentum sums in a fixed order, dead-cell impulses,
//                         sphere step, next band boxes
//   km_retag_band(t+1)    retag inside the band box only
//   km_sort_fresh(t+1)    fresh cells, in the reference's order
//   km_collide(t+1, band) gradients + collide for band cells
//
// Every kernel works on a set of scenes (independent simulations of any size):
// a Scene descriptor in device memory holds a scene's geometry, parameters,
// buffers and buffer parities, and a launch covers the blocks of all scenes,
// each block finding its scene in a prefix table of block offsets.  One scene
// is the single-simulation case; many scenes are the batched form.
//
// The band box holds every cell whose tag can change at t+1, every fresh cell
// and its face neighbours (the only cells whose collide(t+1) reads values the
// fresh pass rewrites), and every cell changed at t.  The sphere moves at most
// kBandDelta cells per axis per step once it is inside its clamp box
// (src/solid.js:159-164, 177-179, 199-205), so a box of radius r + kBandDelta + 2
// around the current centre covers all of them.  Outside the band, tags cannot
// change and no fresh cell is in reach, so collide(t+1) there reads exactly what
// the reference's collide(t+1) reads.
#pragma once

#include "device_common.h"
#include "kernels_pass.h"

namespace lbm {

// Largest per-axis displacement of the sphere centre in one step while it is
// inside its clamp box: |v| <= 0.08 after the speed cap, times 0.98 damping.
constexpr double kBandDelta = 0.09;
constexpr int kBandMargin = 2;

// Axis-aligned band box around a sphere, clamped to the grid.
__host__ __device__ inline void band_box(const lbm_sphere &b, int nx, int ny, int nz, int box[6]) {
  const double c[3] = {b.cx, b.cy, b.cz};
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    const double lo = c[a] - b.r - 0.5 - kBandDelta;
    const double hi = c[a] + b.r - 0.5 + kBandDelta;
    int l, h;
    if (!(lo == lo) || !(hi == hi) || lo < -1e9 || hi > 1e9) {
      l = 0;
      h = n[a] - 1;
    } else {
      l = (int)floor(lo) - kBandMargin;
      h = (int)ceil(hi) + kBandMargin;
    }
    box[2 * a] = l < 0 ? 0 : l;
    box[2 * a + 1] = h > n[a] - 1 ? n[a] - 1 : h;
  }
}

__device__ __forceinline__ bool in_box(const int box[6], int x, int y, int z) {
  return x >= box[0] && x <= box[1] && y >= box[2] && y <= box[3] && z >= box[4] && z <= box[5];
}

// Gradient and Laplacian of a phi array at interior cell i (src/sim.js:186-195),
// rounded to float as the reference's gx/gy/gz/lap arrays store them.
__device__ __forceinline__ void phi_grad(const Geom &g, const float *__restrict__ phi, int64_t i,
                                         double c, float &gxo, float &gyo, float &gzo,
                                         float &lapo) {
  const double px = phi[i + 1], mx = phi[i - 1];
  const double py = phi[i + g.nx], my = phi[i - g.nx];
  const double pz = phi[i + g.nxny], mz = phi[i - g.nxny];
  gxo = (float)(0.5 * (px - mx));
  gyo = (float)(0.5 * (py - my));
  gzo = (float)(0.5 * (pz - mz));
  lapo = (float)(px + mx + py + my + pz + mz - 6 * c);
}

// e_k . v for a lattice direction with components in {-1, 0, 1}.  The reference
// evaluates ekx*vx + eky*vy + ekz*vz; a zero component adds a signed zero, which
// leaves every nonzero sum unchanged, so only the nonzero terms are summed, in
// the same x, y, z order.
template <int EXK, int EYK, int EZK>
__device__ __forceinline__ double edot(double vx, double vy, double vz) {
  double s = 0;
  bool first = true;
  if (EXK != 0) {
    s = EXK > 0 ? vx : -vx;
    first = false;
  }
  if (EYK != 0) {
    const double t = EYK > 0 ? vy : -vy;
    s = first ? t : (EYK > 0 ? s + vy : s - vy);
    first = false;
  }
  if (EZK != 0) {
    const double t = EZK > 0 ? vz : -vz;
    s = first ? t : (EZK > 0 ? s + vz : s - vz);
    first = false;
  }
  return first ? 0.0 : s;
}

// collidePhase for one cell (src/sim.js:341-360).
__device__ __forceinline__ void collide_phase_cell(const Derived &d, double p, double uxi,
                                                   double uyi, double uzi, double ggx, double ggy,
                                                   double ggz, const float hin[Q7],
                                                   float hout[Q7]) {
  const double gmag = sqrt(ggx * ggx + ggy * ggy + ggz * ggz) + 1e-12;
  const double sharpen = d.sharpC * p * (1 - p);
  const double Sx = sharpen * d.sharpFac * ggx / gmag;
  const double Sy = sharpen * d.sharpFac * ggy / gmag;
  const double Sz = sharpen * d.sharpFac * ggz / gmag;
  const double wp0 = (double)w7(0) * p;
  const double wp1 = (double)w7(1) * p;
  const double it = d.invTauPhi;
  // k = 0: eu = 0, source 0.
  {
    const double heq = wp0 * (1 + INV_CS2_7 * 0.0);
    const double hk = hin[0];
    hout[0] = (float)(hk - it * (hk - heq) + (double)w7(0) * 0.0);
  }
  const double u[3] = {uxi, uyi, uzi};
  const double S[3] = {Sx, Sy, Sz};
#pragma unroll
  for (int k = 1; k < Q7; k++) {
    const int a = (k - 1) >> 1;           // axis of direction k
    const bool pos = (k & 1) != 0;        // +1 for odd k
    const double eu = pos ? u[a] : -u[a];
    const double heq = wp1 * (1 + INV_CS2_7 * eu);
    const double Fk = (double)w7(k) * (pos ? S[a] : -S[a]);
    const double hk = hin[k];
    hout[k] = (float)(hk - it * (hk - heq) + Fk);
  }
}

// collideHydro for one cell (src/sim.js:430-459).
template <int K>
__device__ __forceinline__ float collide_hydro_dir(const Derived &d, double sumF, double rhoPhase,
                                                   double negU2half, double uF, double uxe,
                                                   double uye, double uze, double fx, double fy,
                                                   double fz, float fin) {
  constexpr int ex = EX27[K], ey = EY27[K], ez = EZ27[K];
  const double wk = w27(K);
  const double eu = edot<ex, ey, ez>(uxe, uye, uze);
  const double Qi = negU2half + eu * (INV_CS2_27 + INV_2CS4_27 * eu);
  const double feq = wk * (sumF + rhoPhase * Qi);
  const double ekF = edot<ex, ey, ez>(fx, fy, fz);
  const double Sk = d.A * wk * (ekF * (1 + INV_CS2_27 * eu) - uF);
  const double fk = fin;
  return (float)(fk - d.invTau * (fk - feq) + Sk);
}

template <int K>
struct HydroUnroll {
  __device__ __forceinline__ static void run(const Derived &d, double sumF, double rhoPhase,
                                             double negU2half, double uF, double uxe, double uye,
                                             double uze, double fx, double fy, double fz,
                                             const float *fin, float *fout) {
    HydroUnroll<K - 1>::run(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx, fy, fz, fin, fout);
    fout[K - 1] = collide_hydro_dir<K - 1>(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx,
                                           fy, fz, fin[K - 1]);
  }
};
template <>
struct HydroUnroll<0> {
  __device__ __forceinline__ static void run(const Derived &, double, double, double, double,
                                             double, double, double, double, double, double,
                                             const float *, float *) {}
};

__device__ __forceinline__ void collide_hydro_cell(const Derived &d, double sumF, double p,
                                                   double lap, double ggx, double ggy, double ggz,
                                                   double bx, double by, double bz, double uxe,
                                                   double uye, double uze, const float fin[Q27],
                                                   float fout[Q27]) {
  const double rhoPhase = d.rhoG + p * d.dRho;
  const double mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * lap;
  const double fx = bx + mu * ggx;
  const double fy = by + mu * ggy + rhoPhase * d.gravity;
  const double fz = bz + mu * ggz;
  const double u2 = uxe * uxe + uye * uye + uze * uze;
  const double negU2half = -1.5 * u2;
  const double uF = uxe * fx + uye * fy + uze * fz;
  HydroUnroll<Q27>::run(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx, fy, fz, fin, fout);
}

// Pull stream of the 27 populations with halfway bounce-back
// (src/sim.js:509-543).  All source tags are read first and every load address
// is chosen without branching (own opposite population at a non-fluid source),
// so the 27 loads are in flight together; the moving-wall term and the momentum
// exchange run only for solid links.  fk stays the unrounded double the
// reference sums and exchanges; fp receives it rounded to float.
__device__ __forceinline__ void pull27(const Geom &g, int64_t i, const float *__restrict__ f2,
                                       const uint8_t *__restrict__ tag,
                                       const float *__restrict__ usx,
                                       const float *__restrict__ usy,
                                       const float *__restrict__ usz, double rhoPhaseOld,
                                       bool exchange, int x, int y, int z, double cx, double cy,
                                       double cz, double mom[6], bool &linked, float fp[Q27],
                                       double &pSumF, double &mx, double &my, double &mz,
                                       uint32_t &nonfluid) {
  nonfluid = 0;
  uint32_t solid = 0;
#pragma unroll
  for (int k = 1; k < Q27; k++) {
    const int64_t s = i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]);
    const uint8_t st = tag[s];
    nonfluid |= (uint32_t)(st != TAG_FLUID) << k;
    solid |= (uint32_t)(st == TAG_SOLID) << k;
  }
  float raw[Q27];
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    const int64_t s = i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]);
    const int64_t a = ((nonfluid >> k) & 1u) ? opp27(k) * g.S + i : k * g.S + s;
    raw[k] = f2[a];
  }
  pSumF = 0;
  mx = my = mz = 0;
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    double fk = (double)raw[k];
    if ((solid >> k) & 1u) {
      const int ex = EX27[k], ey = EY27[k], ez = EZ27[k];
      const int64_t s = i - ((int64_t)ez * g.nxny + (int64_t)ey * g.nx + ex);
      const double fopp = fk;
      const double eu = (double)ex * (double)usx[s] + (double)ey * (double)usy[s] +
                        (double)ez * (double)usz[s];
      const double corr = 2 * (double)w27(k) * rhoPhaseOld * eu * INV_CS2_27;
      fk = fopp + corr;
      if (exchange) {
        // applyImpulseAtCell(x - ex, y - ey, z - ez, -m ex, -m ey, -m ez)
        const double m = fopp + fk;
        const double px = -m * ex, py = -m * ey, pz = -m * ez;
        const double rx = (x - ex) + 0.5 - cx;
        const double ry = (y - ey) + 0.5 - cy;
        const double rz = (z - ez) + 0.5 - cz;
        mom[0] += px;
        mom[1] += py;
        mom[2] += pz;
        mom[3] += ry * pz - rz * py;
        mom[4] += rz * px - rx * pz;
        mom[5] += rx * py - ry * px;
        linked = true;
      }
    }
    fp[k] = (float)fk;
    pSumF += fk;
    // mx += ekx * fk: a zero component adds a signed zero to a sum that starts at
    // +0, which changes nothing, so only nonzero components are accumulated.
    if (EX27[k] > 0) mx += fk;
    if (EX27[k] < 0) mx += -fk;
    if (EY27[k] > 0) my += fk;
    if (EY27[k] < 0) my += -fk;
    if (EZ27[k] > 0) mz += fk;
    if (EZ27[k] < 0) mz += -fk;
  }
}

// Sum of a double over the 64 lanes of a wave by a fixed butterfly; every lane
// receives the same total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Link momentum of one tile (6 doubles), summed in a fixed order (butterfly
// within each wave, waves in order) and stored at partials[tile].  Every tile
// that intersects the band box stores its sum, zero or not, so the body update
// can read exactly the band's tiles without a list.
__device__ __forceinline__ void tile_momentum(double mom[6], double *partials, int64_t tile) {
  __shared__ double red[4][6];
  const int lane = threadIdx.x;  // 0..63
  const int wave = threadIdx.y;  // 0..3
#pragma unroll
  for (int c = 0; c < 6; c++) mom[c] = wave_sum(mom[c]);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 6; c++) red[wave][c] = mom[c];
  }
  __syncthreads();
  if (lane == 0 && wave == 0) {
#pragma unroll
    for (int c = 0; c < 6; c++)
      partials[tile * 6 + c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
  }
}


// ---------------------------------------------------------------- scenes

// One simulation of a scene set.  Buffer pairs are stored by step parity: a
// step of parity q reads the post-collision populations f[q] (written by the
// previous step) and writes f[q ^ 1]; it reads phi(t) from phi[q] and writes
// phi(t+1) to phi[q ^ 1].  The first collide of a call reads the reference's f
// from f[1] and writes f[0].
struct Scene {
  Geom g;
  TileMap tm;
  Derived d;
  float *f[2], *h[2], *phi[2];
  float *rho, *ux, *uy, *uz;
  uint8_t *tag, *tagPrev;
  float *usx, *usy, *usz;
  float *bfx, *bfy, *bfz;
  int32_t *fresh, *dead;
  int *level_scratch;
  Ctrl *ctrl;
  double *partials;
  double rhoL;              // liquid density, as set (fresh cells use it)
  int list_cap;
  int has_solid;
  int use_bf;               // external force armed for the first step of the call
  int cgx, cgy;             // full-grid blocks of 64x4 cells: cgx * cgy * nz
  int bgx, bgy, bgz;        // band blocks of 64x4 cells (covers the largest band box)
};

// Buffer of parity q (0 or 1) of a pair.  Both pointers are loaded and one is
// selected, so the load does not wait on q.
template <typename T>
__device__ __forceinline__ T *par(T *const pair[2], int q) {
  T *a = pair[0];
  T *b = pair[1];
  return q ? b : a;
}

// Scene of block b: off[s] <= b < off[s + 1].
__device__ __forceinline__ int find_scene(const int *__restrict__ off, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}


// Cell of block lb of a 64x4-block launch over the whole grid or over the band
// box; false when outside.
template <bool BAND>
__device__ __forceinline__ bool block_cell(const Scene &sc, int lb, int &x, int &y, int &z) {
  if (BAND) {
    const int *box = sc.ctrl->box_band;
    x = box[0] + (lb % sc.bgx) * 64 + threadIdx.x;
    y = box[2] + ((lb / sc.bgx) % sc.bgy) * 4 + threadIdx.y;
    z = box[4] + lb / (sc.bgx * sc.bgy);
    return x <= box[1] && y <= box[3] && z <= box[5];
  }
  x = (lb % sc.cgx) * 64 + threadIdx.x;
  y = ((lb / sc.cgx) % sc.cgy) * 4 + threadIdx.y;
  z = lb / (sc.cgx * sc.cgy);
  return x < sc.g.nx && y < sc.g.ny;
}

// Every kernel comes in two launch forms: km_* over a scene table (a block finds
// its scene by binary search in the prefix table off of block offsets) and k1_*
// for one scene passed by value, whose descriptor arrives with the kernel
// arguments and costs no dependent load at block start.
#define LBM_SCENE_KERNELS(NAME, BODY, TPARAMS, TARGS, ...)                               \
  TPARAMS __global__ __launch_bounds__(256) void km_##NAME(                             \
      const Scene *__restrict__ scenes, const int *__restrict__ off, int nscenes,      \
      LBM_PARAMS_##NAME) {                                                               \
    const int si = find_scene(off, nscenes, blockIdx.x);                                 \
    BODY TARGS(scenes[si], blockIdx.x - off[si], __VA_ARGS__);                           \
  }                                                                                      \
  TPARAMS __global__ __launch_bounds__(256) void k1_##NAME(const Scene sc,               \
                                                           LBM_PARAMS_##NAME) {         \
    BODY TARGS(sc, blockIdx.x, __VA_ARGS__);                                             \
  }

// ---------------------------------------------------------------- retag

// retagCells over the whole grid (first step of a call) or over the band box
// (later steps).  No phi snapshot: in the fused form phiPrev is the other phi
// buffer.
template <bool BAND>
__device__ __forceinline__ void retag_body(const Scene &sc, int lb, int) {
  int x, y, z;
  if (!block_cell<BAND>(sc, lb, x, y, z)) return;
  lbm_sphere b{};
  if (sc.has_solid) b = sc.ctrl->sphere;
  retag_cell(sc.g, sc.has_solid ? 1 : 0, b, nullptr, nullptr, sc.tag, sc.tagPrev, nullptr,
             nullptr, 0, sc.usx, sc.usy, sc.usz, sc.ctrl, sc.fresh, sc.dead, sc.list_cap, x, y, z);
}
#define LBM_PARAMS_retag int unused
LBM_SCENE_KERNELS(retag, retag_body, template <bool BAND>, <BAND>, unused)

// Sorts both lists of the scene's last retag, then runs handleFreshCells; one
// workgroup per scene.  Upcoming step p: phi is phi[p]; the post-stream
// populations of fresh cells go to the reference's f (f[1]) before the first
// step of a call, and in place into f[p] for band updates.
__device__ __forceinline__ void sort_fresh_body(const Scene &sc, int p, int prologue) {
  __shared__ int32_t sh[kSortLds];
  int nf = sc.ctrl->n_fresh, nd = sc.ctrl->n_dead;
  if (nf > sc.list_cap) nf = sc.list_cap;
  if (nd > sc.list_cap) nd = sc.list_cap;
  sort_list_block(sc.fresh, nf, sh);
  sort_list_block(sc.dead, nd, sh);
  const int fi = prologue ? 1 : (p & 1);
  float *phi = par(sc.phi, p & 1);
  fresh_block(sc.g, sc.d.rhoG, sc.rhoL, sc.fresh, nf, sc.level_scratch, sc.tag, phi, phi, phi,
              sc.rho, sc.ux, sc.uy, sc.uz, par(sc.f, fi), par(sc.h, fi));
}
__global__ __launch_bounds__(1024) void km_sort_fresh(const Scene *__restrict__ scenes, int p,
                                                      int prologue) {
  sort_fresh_body(scenes[blockIdx.x], p, prologue);
}
__global__ __launch_bounds__(1024) void k1_sort_fresh(const Scene sc, int p, int prologue) {
  sort_fresh_body(sc, p, prologue);
}

// ---------------------------------------------------------------- collide

// Gradients + collide for the fluid cells of the whole grid (first step of a
// call: post-stream populations in f[1] -> post-collision f[0], the external
// force if armed) or of the band box before step p (in place in f[p]).
template <bool BAND>
__device__ __forceinline__ void collide_body(const Scene &sc, int lb, int p) {
  const Geom &g = sc.g;
  int x, y, z;
  if (!block_cell<BAND>(sc, lb, x, y, z)) return;
  if (!interior(g, x, y, z)) return;
  const int64_t i = cell_index(g, x, y, z);
  if (sc.tag[i] != TAG_FLUID) return;
  p &= 1;
  const float *fi = BAND ? par(sc.f, p) : sc.f[1];
  float *fo = BAND ? par(sc.f, p) : sc.f[0];
  const float *hi = BAND ? par(sc.h, p) : sc.h[1];
  float *ho = BAND ? par(sc.h, p) : sc.h[0];
  const float *phi = BAND ? par(sc.phi, p) : sc.phi[0];
  const double pc = phi[i];
  float gx1, gy1, gz1, lap1;
  phi_grad(g, phi, i, pc, gx1, gy1, gz1, lap1);
  const double uxe = sc.ux[i], uye = sc.uy[i], uze = sc.uz[i];
  float hin[Q7], hc[Q7];
#pragma unroll
  for (int k = 0; k < Q7; k++) hin[k] = hi[k * g.S + i];
  collide_phase_cell(sc.d, pc, uxe, uye, uze, (double)gx1, (double)gy1, (double)gz1, hin, hc);
#pragma unroll
  for (int k = 0; k < Q7; k++) ho[k * g.S + i] = hc[k];
  float fin[Q27], fc[Q27];
#pragma unroll
  for (int k = 0; k < Q27; k++) fin[k] = fi[k * g.S + i];
  const bool bf = !BAND && sc.use_bf;
  const double bx = bf ? (double)sc.bfx[i] : 0.0;
  const double by = bf ? (double)sc.bfy[i] : 0.0;
  const double bz = bf ? (double)sc.bfz[i] : 0.0;
  collide_hydro_cell(sc.d, (double)sc.rho[i], pc, (double)lap1, (double)gx1, (double)gy1,
                     (double)gz1, bx, by, bz, uxe, uye, uze, fin, fc);
#pragma unroll
  for (int k = 0; k < Q27; k++) fo[k * g.S + i] = fc[k];
}
#define LBM_PARAMS_collide int p
LBM_SCENE_KERNELS(collide, collide_body, template <bool BAND>, <BAND>, p)

// ---------------------------------------------------------------- phase stream

// phi(t+1) for every cell of step q: fluid cells sum their pulled h
// (src/sim.js:376-397, 549-552); other cells keep phi (the reference never
// updates them).  Source tags are read first and every load address is chosen
// without branching, so the seven loads are in flight together.
__device__ __forceinline__ void phase_stream_body(const Scene &sc, int lb, int q) {
  const Geom &g = sc.g;
  int tile, z, c0, x, y;
  if (!map_tile(sc.tm, lb, tile, z, c0)) return;
  if (!tile_cell(sc.tm, c0, x, y)) return;
  q &= 1;
  const float *h2 = par(sc.h, q);
  const float *phi = par(sc.phi, q);
  float *phi_next = par(sc.phi, q ^ 1);
  const uint8_t *tag = sc.tag;
  const int64_t i = cell_index(g, x, y, z);
  if (tag[i] != TAG_FLUID || !interior(g, x, y, z)) {
    phi_next[i] = phi[i];
    return;
  }
  uint32_t nonfluid = 0;
#pragma unroll
  for (int k = 1; k < Q7; k++) {
    const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
    nonfluid |= (uint32_t)(tag[s] != TAG_FLUID) << k;
  }
  float hv[Q7];
#pragma unroll
  for (int k = 0; k < Q7; k++) {
    const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
    const int64_t a = ((nonfluid >> k) & 1u) ? opp7(k) * g.S + i : k * g.S + s;
    hv[k] = h2[a];
  }
  double phiSum = 0;
#pragma unroll
  for (int k = 0; k < Q7; k++) phiSum += (double)hv[k];
  const double pp = phiSum < 0 ? 0 : phiSum > 1 ? 1 : phiSum;
  phi_next[i] = (float)pp;
}
#define LBM_PARAMS_phase_stream int q
LBM_SCENE_KERNELS(phase_stream, phase_stream_body, , , q)

// ---------------------------------------------------------------- stream + collide

// stream(t) + macro(t) for every fluid cell of step q; with COLLIDE, also
// collide(t+1) for fluid cells outside the band.  Without COLLIDE (the last
// step of a call) the post-stream populations are stored for every cell, which
// leaves the state exactly as the reference's step does, and phi(t+1) is
// written here.  first: the first step of the call (the armed external force
// applies to its macroscopic update).
template <bool COLLIDE>
__device__ __forceinline__ void stream_collide_body(const Scene &sc, int lb, int q, int first) {
  const Geom &g = sc.g;
  const Derived &d = sc.d;
  int tile, z, c0, x, y;
  if (!map_tile(sc.tm, lb, tile, z, c0)) return;
  const bool in_plane = tile_cell(sc.tm, c0, x, y);
  q &= 1;
  const float *__restrict__ f2 = par(sc.f, q);
  const float *__restrict__ h2 = par(sc.h, q);
  float *__restrict__ fo = par(sc.f, q ^ 1);
  float *__restrict__ ho = par(sc.h, q ^ 1);
  const float *__restrict__ phi = par(sc.phi, q);
  float *__restrict__ phi_next = par(sc.phi, q ^ 1);
  const uint8_t *__restrict__ tag = sc.tag;
  const Ctrl *__restrict__ ctrl = sc.ctrl;
  double mom[6] = {0, 0, 0, 0, 0, 0};
  bool linked = false;
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = ctrl->box_fused[j];
  // Link momentum can only arise next to the sphere, which lies inside the band:
  // the tiles of the band's planes between its first and last cell store their
  // sums (zero or not), and the body update reads exactly those.
  int tlo, thi;
  box_tiles(sc.tm, box, tlo, thi);
  const int lt = tile - z * sc.tm.tpp;
  const bool tile_in_band = z >= box[4] && z <= box[5] && lt >= tlo && lt <= thi;
  const bool exchange = sc.has_solid != 0 && tile_in_band;
  const int use_bf = first && sc.use_bf;
  double cx = 0, cy = 0, cz = 0;
  if (exchange) {
    cx = ctrl->sphere.cx;
    cy = ctrl->sphere.cy;
    cz = ctrl->sphere.cz;
  }
  if (in_plane) {
    const int64_t i = cell_index(g, x, y, z);
    const bool fluid = tag[i] == TAG_FLUID && interior(g, x, y, z);
    if (!fluid) {
      if (!COLLIDE) phi_next[i] = phi[i];
    } else {
      const double phiOld = phi[i];
      const double rhoPhaseOld = d.rhoG + phiOld * d.dRho;
      float gxo, gyo, gzo, lapo;
      phi_grad(g, phi, i, phiOld, gxo, gyo, gzo, lapo);
      float fp[Q27];
      double pSumF, mx, my, mz;
      uint32_t nonfluid;
      pull27(g, i, f2, tag, sc.usx, sc.usy, sc.usz, rhoPhaseOld, exchange, x, y, z, cx, cy, cz,
             mom, linked, fp, pSumF, mx, my, mz, nonfluid);
      const double rhoN = pSumF > 0.01 ? pSumF : 0.01;
      // D3Q7 directions 1..6 are D3Q27 directions 1..6, so the same source mask
      // selects their bounce-back.
      float hp[Q7];
#pragma unroll
      for (int k = 0; k < Q7; k++) {
        const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
        const int64_t a = ((nonfluid >> k) & 1u) ? opp7(k) * g.S + i : k * g.S + s;
        hp[k] = h2[a];
      }
      double phiSum = 0;
#pragma unroll
      for (int k = 0; k < Q7; k++) phiSum += (double)hp[k];
      const double p = phiSum < 0 ? 0 : phiSum > 1 ? 1 : phiSum;
      const double rhoPhaseNew = d.rhoG + p * d.dRho;
      const double mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * (double)lapo;
      const double bxf = use_bf ? (double)sc.bfx[i] : 0.0;
      const double byf = use_bf ? (double)sc.bfy[i] : 0.0;
      const double bzf = use_bf ? (double)sc.bfz[i] : 0.0;
      const double fx = bxf + mu * (double)gxo;
      const double fy = byf + mu * (double)gyo + rhoPhaseNew * d.gravity;
      const double fz = bzf + mu * (double)gzo;
      const double invRho = 1 / rhoPhaseNew;
      double uxn = (mx + 0.5 * fx) * invRho;
      double uyn = (my + 0.5 * fy) * invRho;
      double uzn = (mz + 0.5 * fz) * invRho;
      const double u2 = uxn * uxn + uyn * uyn + uzn * uzn;
      const double UMAX = 0.2, UMAX2 = UMAX * UMAX;
      if (u2 > UMAX2) {
        const double scl = UMAX / sqrt(u2);
        uxn *= scl;
        uyn *= scl;
        uzn *= scl;
      }
      const float rhof = (float)rhoN, phif = (float)p;
      const float uxf = (float)uxn, uyf = (float)uyn, uzf = (float)uzn;
      const bool band = in_box(box, x, y, z);
      // rho and u are read before the end of the call only inside the band (by
      // the fresh-cell pass, the dead-cell impulses and the band collide); the
      // last step of a call stores them everywhere.
      if (!COLLIDE || band) {
        sc.rho[i] = rhof;
        sc.ux[i] = uxf;
        sc.uy[i] = uyf;
        sc.uz[i] = uzf;
      }
      bool store_stream = true;
      if (COLLIDE) {
        if (!band) {
          store_stream = false;
          // collide(t+1) with the new macroscopic values (rounded to float as
          // the reference stores them) and the gradients of phi(t+1).
          float gx1, gy1, gz1, lap1;
          phi_grad(g, phi_next, i, (double)phif, gx1, gy1, gz1, lap1);
          float hc[Q7];
          collide_phase_cell(d, (double)phif, (double)uxf, (double)uyf, (double)uzf, (double)gx1,
                             (double)gy1, (double)gz1, hp, hc);
#pragma unroll
          for (int k = 0; k < Q7; k++) ho[k * g.S + i] = hc[k];
          float fc[Q27];
          collide_hydro_cell(d, (double)rhof, (double)phif, (double)lap1, (double)gx1,
                             (double)gy1, (double)gz1, 0.0, 0.0, 0.0, (double)uxf, (double)uyf,
                             (double)uzf, fp, fc);
#pragma unroll
          for (int k = 0; k < Q27; k++) fo[k * g.S + i] = fc[k];
        }
      } else {
        phi_next[i] = phif;
      }
      if (store_stream) {
#pragma unroll
        for (int k = 0; k < Q27; k++) fo[k * g.S + i] = fp[k];
#pragma unroll
        for (int k = 0; k < Q7; k++) ho[k * g.S + i] = hp[k];
      }
    }
  }
  (void)linked;
  if (exchange) tile_momentum(mom, sc.partials, tile);
}
#define LBM_PARAMS_stream_collide int q, int first
LBM_SCENE_KERNELS(stream_collide, stream_collide_body, template <bool COLLIDE>, <COLLIDE>, q,
                  first)

// ---------------------------------------------------------------- body

// Body update after stream(t) of step q, one 512-thread workgroup per scene:
//   F/T += link momentum of the band's tiles (in each plane of this step's
//          band box, the tiles from its first to its last cell), in ascending
//          tile order (contiguous chunks per thread, then waves in order);
//   F/T += dead-cell impulses rho_phase(phi) u at the cells that turned solid at
//          this step's retag (src/solid.js:87-101), in the same chunked order;
//          a dead cell keeps the phi and u it had when it turned solid, which
//          phi(t) holds;
//   RigidSphere.step(1, g) (src/solid.js:104-210).
// Then rolls the band boxes (the box of the old position serves the band
// kernels of the next step, the box of the new position the next fused
// launch) and, when check_band is set, verifies the centre moved at most
// kBandDelta per axis.
constexpr int kBodyThreads = 512;
__device__ __forceinline__ void body_body(const Scene &sc, int q, int check_band,
                                          int reset_lists) {
  __shared__ double red[kBodyThreads / 64][12];
  const Geom &g = sc.g;
  const TileMap &tm = sc.tm;
  Ctrl *ctrl = sc.ctrl;
  const int tid = threadIdx.x;
  if (!sc.has_solid) {
    if (tid == 0 && reset_lists) {
      ctrl->n_fresh = 0;
      ctrl->n_dead = 0;
      ctrl->n_solid = 0;
    }
    return;
  }
  const double c0x = ctrl->sphere.cx, c0y = ctrl->sphere.cy, c0z = ctrl->sphere.cz;
  const float *phi = par(sc.phi, q & 1);
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = ctrl->box_fused[j];
  int tlo, thi;
  box_tiles(tm, box, tlo, thi);
  const int nbt = thi - tlo + 1;
  const int bz0 = box[4], nbz = box[5] - box[4] + 1;
  const int ntiles = nbt * nbz;
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = 0;
  {
    const int C = (ntiles + kBodyThreads - 1) / kBodyThreads;
    const int lo = tid * C, hi = min(ntiles, lo + C);
    for (int c = lo; c < hi; c++) {
      const int tz_ = c / nbt;
      const int64_t t = (int64_t)(bz0 + tz_) * tm.tpp + tlo + (c - tz_ * nbt);
#pragma unroll
      for (int k = 0; k < 6; k++) acc[k] += sc.partials[t * 6 + k];
    }
  }
  int nd = ctrl->n_dead;
  if (nd > sc.list_cap) nd = sc.list_cap;
  {
    const int C = (nd + kBodyThreads - 1) / kBodyThreads;
    const int lo = tid * C, hi = min(nd, lo + C);
    for (int c = lo; c < hi; c++) {
      const int32_t i = sc.dead[c];
      const int x = (int)(i % g.nx), y = (int)((i / g.nx) % g.ny), z = (int)(i / g.nxny);
      const double rhoPhase = sc.d.rhoG + (double)phi[i] * sc.d.dRho;
      const double px = rhoPhase * (double)sc.ux[i];
      const double py = rhoPhase * (double)sc.uy[i];
      const double pz = rhoPhase * (double)sc.uz[i];
      const double rx = x + 0.5 - c0x;
      const double ry = y + 0.5 - c0y;
      const double rz = z + 0.5 - c0z;
      acc[6] += px;
      acc[7] += py;
      acc[8] += pz;
      acc[9] += ry * pz - rz * py;
      acc[10] += rz * px - rx * pz;
      acc[11] += rx * py - ry * px;
    }
  }
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (tid != 0) return;
  double tot[12];
  for (int k = 0; k < 12; k++) {
    double v = 0;
    for (int w = 0; w < kBodyThreads / 64; w++) v += red[w][k];
    tot[k] = v;
  }
  lbm_sphere b = ctrl->sphere;
  b.Fx += tot[0];
  b.Fy += tot[1];
  b.Fz += tot[2];
  b.Tx += tot[3];
  b.Ty += tot[4];
  b.Tz += tot[5];
  for (int k = 0; k < 6; k++) ctrl->F_links[k] = tot[k];
  b.Fx += tot[6];
  b.Fy += tot[7];
  b.Fz += tot[8];
  b.Tx += tot[9];
  b.Ty += tot[10];
  b.Tz += tot[11];
  sphere_step(b, 1.0, sc.d.gravity);
  if (check_band) {
    const double d0 = b.cx - c0x, d1 = b.cy - c0y, d2 = b.cz - c0z;
    if (!(d0 <= kBandDelta && d0 >= -kBandDelta && d1 <= kBandDelta && d1 >= -kBandDelta &&
          d2 <= kBandDelta && d2 >= -kBandDelta))
      ctrl->err |= kErrBandEscape;
  }
  ctrl->sphere = b;
  for (int j = 0; j < 6; j++) ctrl->box_band[j] = box[j];
  int nb[6];
  band_box(b, g.nx, g.ny, g.nz, nb);
  for (int j = 0; j < 6; j++) ctrl->box_fused[j] = nb[j];
  if (reset_lists) {
    ctrl->n_fresh = 0;
    ctrl->n_dead = 0;
    ctrl->n_solid = 0;
  }
}
__global__ __launch_bounds__(kBodyThreads) void km_body(const Scene *__restrict__ scenes, int q,
                                                        int check_band, int reset_lists) {
  body_body(scenes[blockIdx.x], q, check_band, reset_lists);
}
__global__ __launch_bounds__(kBodyThreads) void k1_body(const Scene sc, int q, int check_band,
                                                        int reset_lists) {
  body_body(sc, q, check_band, reset_lists);
}

}  // namespace lbm
