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

#include "codec.h"
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
// rounded to float as the reference's gx/gy/gz/lap arrays store them.  R is the
// arithmetic type (double for the reference-exact path).
template <typename R>
__device__ __forceinline__ void phi_grad(const Geom &g, const float *__restrict__ phi, int64_t i,
                                         R c, float &gxo, float &gyo, float &gzo, float &lapo) {
  const R px = phi[i + 1], mx = phi[i - 1];
  const R py = phi[i + g.nx], my = phi[i - g.nx];
  const R pz = phi[i + g.nxny], mz = phi[i - g.nxny];
  gxo = (float)(R(0.5) * (px - mx));
  gyo = (float)(R(0.5) * (py - my));
  gzo = (float)(R(0.5) * (pz - mz));
  lapo = (float)(px + mx + py + my + pz + mz - 6 * c);
}

// e_k . v for a lattice direction with components in {-1, 0, 1}.  The reference
// evaluates ekx*vx + eky*vy + ekz*vz; a zero component adds a signed zero, which
// leaves every nonzero sum unchanged, so only the nonzero terms are summed, in
// the same x, y, z order.
template <int EXK, int EYK, int EZK, typename R>
__device__ __forceinline__ R edot(R vx, R vy, R vz) {
  R s = 0;
  bool first = true;
  if (EXK != 0) {
    s = EXK > 0 ? vx : -vx;
    first = false;
  }
  if (EYK != 0) {
    const R t = EYK > 0 ? vy : -vy;
    s = first ? t : (EYK > 0 ? s + vy : s - vy);
    first = false;
  }
  if (EZK != 0) {
    const R t = EZK > 0 ? vz : -vz;
    s = first ? t : (EZK > 0 ? s + vz : s - vz);
    first = false;
  }
  return first ? R(0) : s;
}

// collidePhase for one cell (src/sim.js:341-360).
template <typename R>
__device__ __forceinline__ void collide_phase_cell(const Derived &d, R p, R uxi, R uyi, R uzi,
                                                   R ggx, R ggy, R ggz, const float hin[Q7],
                                                   float hout[Q7]) {
  const R gmag = sqrt(ggx * ggx + ggy * ggy + ggz * ggz) + R(1e-12);
  const R sharpen = R(d.sharpC) * p * (1 - p);
  const R Sx = sharpen * R(d.sharpFac) * ggx / gmag;
  const R Sy = sharpen * R(d.sharpFac) * ggy / gmag;
  const R Sz = sharpen * R(d.sharpFac) * ggz / gmag;
  const R wp0 = (R)w7(0) * p;
  const R wp1 = (R)w7(1) * p;
  const R it = R(d.invTauPhi);
  const R c7 = R(INV_CS2_7);
  // k = 0: eu = 0, source 0.
  {
    const R heq = wp0 * (1 + c7 * R(0));
    const R hk = hin[0];
    hout[0] = (float)(hk - it * (hk - heq) + (R)w7(0) * R(0));
  }
  const R u[3] = {uxi, uyi, uzi};
  const R S[3] = {Sx, Sy, Sz};
#pragma unroll
  for (int k = 1; k < Q7; k++) {
    const int a = (k - 1) >> 1;           // axis of direction k
    const bool pos = (k & 1) != 0;        // +1 for odd k
    const R eu = pos ? u[a] : -u[a];
    const R heq = wp1 * (1 + c7 * eu);
    const R Fk = (R)w7(k) * (pos ? S[a] : -S[a]);
    const R hk = hin[k];
    hout[k] = (float)(hk - it * (hk - heq) + Fk);
  }
}

// collideHydro for one cell (src/sim.js:430-459).
template <int K, typename R>
__device__ __forceinline__ float collide_hydro_dir(const Derived &d, R sumF, R rhoPhase,
                                                   R negU2half, R uF, R uxe, R uye, R uze, R fx,
                                                   R fy, R fz, float fin) {
  constexpr int ex = EX27[K], ey = EY27[K], ez = EZ27[K];
  const R wk = (R)w27(K);
  const R eu = edot<ex, ey, ez>(uxe, uye, uze);
  const R Qi = negU2half + eu * (R(INV_CS2_27) + R(INV_2CS4_27) * eu);
  const R feq = wk * (sumF + rhoPhase * Qi);
  const R ekF = edot<ex, ey, ez>(fx, fy, fz);
  const R Sk = R(d.A) * wk * (ekF * (1 + R(INV_CS2_27) * eu) - uF);
  const R fk = fin;
  return (float)(fk - R(d.invTau) * (fk - feq) + Sk);
}

template <int K, typename R>
struct HydroUnroll {
  __device__ __forceinline__ static void run(const Derived &d, R sumF, R rhoPhase, R negU2half,
                                             R uF, R uxe, R uye, R uze, R fx, R fy, R fz,
                                             const float *fin, float *fout) {
    HydroUnroll<K - 1, R>::run(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx, fy, fz, fin,
                               fout);
    fout[K - 1] = collide_hydro_dir<K - 1, R>(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx,
                                              fy, fz, fin[K - 1]);
  }
};
template <typename R>
struct HydroUnroll<0, R> {
  __device__ __forceinline__ static void run(const Derived &, R, R, R, R, R, R, R, R, R, R,
                                             const float *, float *) {}
};

template <typename R>
__device__ __forceinline__ void collide_hydro_cell(const Derived &d, R sumF, R p, R lap, R ggx,
                                                   R ggy, R ggz, R bx, R by, R bz, R uxe, R uye,
                                                   R uze, const float fin[Q27], float fout[Q27]) {
  const R rhoPhase = R(d.rhoG) + p * R(d.dRho);
  const R mu = 4 * R(d.beta) * p * (p - 1) * (2 * p - 1) - R(d.kappa) * lap;
  const R fx = bx + mu * ggx;
  const R fy = by + mu * ggy + rhoPhase * R(d.gravity);
  const R fz = bz + mu * ggz;
  const R u2 = uxe * uxe + uye * uye + uze * uze;
  const R negU2half = R(-1.5) * u2;
  const R uF = uxe * fx + uye * fy + uze * fz;
  HydroUnroll<Q27, R>::run(d, sumF, rhoPhase, negU2half, uF, uxe, uye, uze, fx, fy, fz, fin, fout);
}

// Direction masks: bit k set when the component of e_k along the axis is
// positive (P) or negative (M).
template <int A, int SIGN>
struct DirMask {
  static constexpr uint32_t make() {
    uint32_t m = 0;
    for (int k = 0; k < Q27; k++) {
      const int e = A == 0 ? EX27[k] : A == 1 ? EY27[k] : EZ27[k];
      if (e == SIGN) m |= 1u << k;
    }
    return m;
  }
  static constexpr uint32_t v = make();
};

// Source mask of interior cell (x, y, z) when no solid cell is within reach:
// walls are exactly the grid faces (src/sim.js:226-229), so source x - e_x is a
// wall when x = 1 and e_x = +1 or x = nx - 2 and e_x = -1, and likewise for y
// and z.  Bit k is set when the source of direction k is not fluid.
__device__ __forceinline__ uint32_t wall_mask(const Geom &g, int x, int y, int z) {
  uint32_t m = 0;
  if (x == 1) m |= DirMask<0, 1>::v;
  if (x == g.nx - 2) m |= DirMask<0, -1>::v;
  if (y == 1) m |= DirMask<1, 1>::v;
  if (y == g.ny - 2) m |= DirMask<1, -1>::v;
  if (z == 1) m |= DirMask<2, 1>::v;
  if (z == g.nz - 2) m |= DirMask<2, -1>::v;
  return m;
}

// Source masks from the neighbours' tags (bit k: source of direction k not
// fluid / solid).
__device__ __forceinline__ void tag_masks(const Geom &g, int64_t i,
                                          const uint8_t *__restrict__ tag, uint32_t &nonfluid,
                                          uint32_t &solid) {
  nonfluid = 0;
  solid = 0;
#pragma unroll
  for (int k = 1; k < Q27; k++) {
    const int64_t s = i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]);
    const uint8_t st = tag[s];
    nonfluid |= (uint32_t)(st != TAG_FLUID) << k;
    solid |= (uint32_t)(st == TAG_SOLID) << k;
  }
}

// Pull stream of the 27 populations with halfway bounce-back
// (src/sim.js:509-543), given the source masks.  Every load address is chosen
// without branching (own opposite population at a non-fluid source), so the 27
// loads are in flight together; the moving-wall term and the momentum
// exchange run only for solid links.  fk stays the unrounded value the
// reference sums and exchanges; fp receives it rounded to float.  The momentum
// exchange accumulates in FP64 whatever the arithmetic type.
template <int KIND>
__device__ __forceinline__ void pull27(const Geom &g, int64_t i,
                                       const typename Codec<KIND>::T *__restrict__ f2,
                                       uint32_t nonfluid, uint32_t solid,
                                       const float *__restrict__ usx,
                                       const float *__restrict__ usy,
                                       const float *__restrict__ usz,
                                       typename Codec<KIND>::R rhoPhaseOld, bool exchange, int x,
                                       int y, int z, double cx, double cy, double cz,
                                       double mom[6], float fp[Q27],
                                       typename Codec<KIND>::R &pSumF,
                                       typename Codec<KIND>::R &mx, typename Codec<KIND>::R &my,
                                       typename Codec<KIND>::R &mz) {
  using C = Codec<KIND>;
  using R = typename C::R;
  float raw[Q27];
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    const int64_t s = i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]);
    const int64_t a = ((nonfluid >> k) & 1u) ? opp27(k) * g.S + i : k * g.S + s;
    raw[k] = C::ld(f2, a, w27(k));  // w27(opp(k)) == w27(k)
  }
  pSumF = 0;
  mx = my = mz = 0;
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    R fk = (R)raw[k];
    if ((solid >> k) & 1u) {
      const int ex = EX27[k], ey = EY27[k], ez = EZ27[k];
      const int64_t s = i - ((int64_t)ez * g.nxny + (int64_t)ey * g.nx + ex);
      const R fopp = fk;
      const R eu = (R)ex * (R)usx[s] + (R)ey * (R)usy[s] + (R)ez * (R)usz[s];
      const R corr = 2 * (R)w27(k) * rhoPhaseOld * eu * R(INV_CS2_27);
      fk = fopp + corr;
      if (exchange) {
        // applyImpulseAtCell(x - ex, y - ey, z - ez, -m ex, -m ey, -m ez)
        const double m = (double)fopp + (double)fk;
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

// Pull of the 7 phase populations (src/sim.js:376-397): from the source cell,
// or the own opposite population when the source is not fluid (D3Q7 directions
// 1..6 are D3Q27 directions 1..6, so the D3Q27 source mask serves).  A stored h
// carries the phi of the cell it is stored at as its baseline.
template <int KIND>
__device__ __forceinline__ void pull7(const Geom &g, int64_t i,
                                      const typename Codec<KIND>::T *__restrict__ h2,
                                      const float *__restrict__ phi, uint32_t nonfluid,
                                      float hp[Q7]) {
  using C = Codec<KIND>;
#pragma unroll
  for (int k = 0; k < Q7; k++) {
    const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
    const bool bounce = (nonfluid >> k) & 1u;
    const int64_t a = bounce ? opp7(k) * g.S + i : k * g.S + s;
    float base = 0.0f;
    if (KIND != ST_F32) base = w7(k) * phi[bounce ? i : s];
    hp[k] = C::ld(h2, a, base);
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
// of the band stores its sum, zero or not, so the body update can read exactly
// the band's tiles without a list.
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
// from f[1] and writes f[0].  The population pointers address elements of the
// scene set's storage codec (float, or 16-bit deviations).
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
  int live;                 // continues a call that ended after a step's body update
  int use_bf;               // external force armed for the first step of the call
  int cgx, cgy;             // full-grid blocks of 64x4 cells: cgx * cgy * nz
  int bgx, bgy, bgz;        // band blocks of 64x4 cells (covers the largest band box)
};

// Buffer of parity q (0 or 1) of a pair, as codec elements.  Both pointers are
// loaded and one is selected, so the load does not wait on q.
template <int KIND>
__device__ __forceinline__ typename Codec<KIND>::T *pop(float *const pair[2], int q) {
  float *a = pair[0];
  float *b = pair[1];
  return reinterpret_cast<typename Codec<KIND>::T *>(q ? b : a);
}
__device__ __forceinline__ float *par(float *const pair[2], int q) {
  float *a = pair[0];
  float *b = pair[1];
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
// arguments and costs no dependent load at block start.  KIND selects the
// population codec of every scene of the launch.

// ---------------------------------------------------------------- retag

// retagCells over the whole grid (first step of a call) or over the band box
// (later steps).  No phi snapshot: in the fused form phiPrev is the other phi
// buffer.
template <bool BAND>
__device__ __forceinline__ void retag_body(const Scene &sc, int lb) {
  int x, y, z;
  if (!block_cell<BAND>(sc, lb, x, y, z)) return;
  lbm_sphere b{};
  if (sc.has_solid) b = sc.ctrl->sphere;
  retag_cell(sc.g, sc.has_solid ? 1 : 0, b, nullptr, nullptr, sc.tag, sc.tagPrev, nullptr,
             nullptr, 0, sc.usx, sc.usy, sc.usz, sc.ctrl, sc.fresh, sc.dead, sc.list_cap, x, y, z);
}
template <bool BAND>
__global__ __launch_bounds__(256) void km_retag(const Scene *__restrict__ scenes,
                                                const int *__restrict__ off, int nscenes) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  retag_body<BAND>(scenes[si], blockIdx.x - off[si]);
}
template <bool BAND>
__global__ __launch_bounds__(256) void k1_retag(const Scene sc) {
  retag_body<BAND>(sc, blockIdx.x);
}

// Sorts both lists of the scene's last retag, then runs handleFreshCells; one
// workgroup per scene.  Upcoming step p: phi is phi[p]; the post-stream
// populations of fresh cells go to the reference's f (f[1]) before the first
// step of a call (mode 0, scenes that start from the reference's state), and
// in place into f[p] for band updates (mode 1, every scene; mode 2, the band
// update a continued scene owes before its first step).
template <int KIND>
__device__ __forceinline__ void sort_fresh_body(const Scene &sc, int p, int mode) {
  if ((mode == 0 && sc.live) || (mode == 2 && !sc.live)) return;
  const int prologue = mode == 0;
  __shared__ int32_t sh[kSortLds];
  int nf = sc.ctrl->n_fresh, nd = sc.ctrl->n_dead;
  if (nf > sc.list_cap) nf = sc.list_cap;
  if (nd > sc.list_cap) nd = sc.list_cap;
  sort_list_block(sc.fresh, nf, sh);
  sort_list_block(sc.dead, nd, sh);
  const int fi = prologue ? 1 : (p & 1);
  float *phi = par(sc.phi, p & 1);
  fresh_block<KIND>(sc.g, sc.d.rhoG, sc.rhoL, sc.fresh, nf, sc.level_scratch, sc.tag, phi, phi,
                    phi, sc.rho, sc.ux, sc.uy, sc.uz, pop<KIND>(sc.f, fi), pop<KIND>(sc.h, fi));
}
template <int KIND>
__global__ __launch_bounds__(1024) void km_sort_fresh(const Scene *__restrict__ scenes, int p,
                                                      int mode) {
  sort_fresh_body<KIND>(scenes[blockIdx.x], p, mode);
}
template <int KIND>
__global__ __launch_bounds__(1024) void k1_sort_fresh(const Scene sc, int p, int mode) {
  sort_fresh_body<KIND>(sc, p, mode);
}

// ---------------------------------------------------------------- collide

// Gradients + collide for the fluid cells of the whole grid (first step of a
// call: post-stream populations in f[1] -> post-collision f[0], the external
// force if armed) or of the band box before step p (in place in f[p]).  The h
// baseline is phi of the same cell before and after: the collide keeps phi.
template <bool BAND, int KIND>
__device__ __forceinline__ void collide_body(const Scene &sc, int lb, int p) {
  using C = Codec<KIND>;
  using R = typename C::R;
  using T = typename C::T;
  const Geom &g = sc.g;
  int x, y, z;
  if (!block_cell<BAND>(sc, lb, x, y, z)) return;
  if (!interior(g, x, y, z)) return;
  const int64_t i = cell_index(g, x, y, z);
  if (sc.tag[i] != TAG_FLUID) return;
  p &= 1;
  const T *fi = BAND ? pop<KIND>(sc.f, p) : pop<KIND>(sc.f, 1);
  T *fo = BAND ? pop<KIND>(sc.f, p) : pop<KIND>(sc.f, 0);
  const T *hi = BAND ? pop<KIND>(sc.h, p) : pop<KIND>(sc.h, 1);
  T *ho = BAND ? pop<KIND>(sc.h, p) : pop<KIND>(sc.h, 0);
  const float *phi = BAND ? par(sc.phi, p) : sc.phi[0];
  const float phc = phi[i];
  const R pc = phc;
  float gx1, gy1, gz1, lap1;
  phi_grad<R>(g, phi, i, pc, gx1, gy1, gz1, lap1);
  const R uxe = sc.ux[i], uye = sc.uy[i], uze = sc.uz[i];
  float hin[Q7], hc[Q7];
#pragma unroll
  for (int k = 0; k < Q7; k++) hin[k] = C::ld(hi, k * g.S + i, w7(k) * phc);
  collide_phase_cell<R>(sc.d, pc, uxe, uye, uze, (R)gx1, (R)gy1, (R)gz1, hin, hc);
#pragma unroll
  for (int k = 0; k < Q7; k++) C::st(ho, k * g.S + i, hc[k], w7(k) * phc);
  float fin[Q27], fc[Q27];
#pragma unroll
  for (int k = 0; k < Q27; k++) fin[k] = C::ld(fi, k * g.S + i, w27(k));
  const bool bf = !BAND && sc.use_bf;
  const R bx = bf ? (R)sc.bfx[i] : R(0);
  const R by = bf ? (R)sc.bfy[i] : R(0);
  const R bz = bf ? (R)sc.bfz[i] : R(0);
  collide_hydro_cell<R>(sc.d, (R)sc.rho[i], pc, (R)lap1, (R)gx1, (R)gy1, (R)gz1, bx, by, bz, uxe,
                        uye, uze, fin, fc);
#pragma unroll
  for (int k = 0; k < Q27; k++) C::st(fo, k * g.S + i, fc[k], w27(k));
}
template <bool BAND, int KIND>
__global__ __launch_bounds__(256) void km_collide(const Scene *__restrict__ scenes,
                                                  const int *__restrict__ off, int nscenes,
                                                  int p) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  collide_body<BAND, KIND>(scenes[si], blockIdx.x - off[si], p);
}
template <bool BAND, int KIND>
__global__ __launch_bounds__(256) void k1_collide(const Scene sc, int p) {
  collide_body<BAND, KIND>(sc, blockIdx.x, p);
}

// ---------------------------------------------------------------- phase stream

// phi(t+1) for every cell of step q: fluid cells sum their pulled h
// (src/sim.js:376-397, 549-552); other cells keep phi (the reference never
// updates them).  Source tags are read first and every load address is chosen
// without branching, so the seven loads are in flight together.
template <int KIND>
__device__ __forceinline__ void phase_stream_body(const Scene &sc, int lb, int q) {
  using R = typename Codec<KIND>::R;
  const Geom &g = sc.g;
  int tile, x, y, z;
  if (!map_tile(sc.tm.T, sc.tm.T8, lb, tile)) return;
  const int64_t i = ((int64_t)tile << kTileShift) + threadIdx.y * 64 + threadIdx.x;
  if (i >= g.n) return;
  cell_xyz(g.nx, g.nxny, i, x, y, z);
  q &= 1;
  const typename Codec<KIND>::T *h2 = pop<KIND>(sc.h, q);
  const float *phi = par(sc.phi, q);
  float *phi_next = par(sc.phi, q ^ 1);
  const uint8_t *tag = sc.tag;
  // Outside the band box the only non-fluid cells are the walls on the grid
  // faces (solid cells lie inside it with a margin of 2).
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = sc.ctrl->box_fused[j];
  const bool near = in_box(box, x, y, z);
  if (!interior(g, x, y, z) || (near && tag[i] != TAG_FLUID)) {
    phi_next[i] = phi[i];
    return;
  }
  uint32_t nonfluid = 0;
  if (near) {
#pragma unroll
    for (int k = 1; k < Q7; k++) {
      const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
      nonfluid |= (uint32_t)(tag[s] != TAG_FLUID) << k;
    }
  } else {
    nonfluid = wall_mask(g, x, y, z) & 0x7eu;  // D3Q7 directions 1..6
  }
  float hv[Q7];
  pull7<KIND>(g, i, h2, phi, nonfluid, hv);
  R phiSum = 0;
#pragma unroll
  for (int k = 0; k < Q7; k++) phiSum += (R)hv[k];
  const R pp = phiSum < 0 ? R(0) : phiSum > 1 ? R(1) : phiSum;
  phi_next[i] = (float)pp;
}
template <int KIND>
__global__ __launch_bounds__(256) void km_phase_stream(const Scene *__restrict__ scenes,
                                                       const int *__restrict__ off, int nscenes,
                                                       int q) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  phase_stream_body<KIND>(scenes[si], blockIdx.x - off[si], q);
}
template <int KIND>
__global__ __launch_bounds__(256) void k1_phase_stream(const Scene sc, int q) {
  phase_stream_body<KIND>(sc, blockIdx.x, q);
}

// ---------------------------------------------------------------- stream + collide

// stream(t) + macro(t) for every fluid cell of step q; with COLLIDE, also
// collide(t+1) for fluid cells outside the band.  Without COLLIDE (the last
// step of a call) the post-stream populations are stored for every cell, which
// leaves the state exactly as the reference's step does, and phi(t+1) is
// written here.  first: the first step of the call (the armed external force
// applies to its macroscopic update).
template <bool COLLIDE, int KIND>
__device__ __forceinline__ void stream_collide_body(const Scene &sc, int lb, int q, int first) {
  using C = Codec<KIND>;
  using R = typename C::R;
  using T = typename C::T;
  const Geom &g = sc.g;
  const Derived &d = sc.d;
  int tile, x, y, z;
  if (!map_tile(sc.tm.T, sc.tm.T8, lb, tile)) return;
  const int64_t i = ((int64_t)tile << kTileShift) + threadIdx.y * 64 + threadIdx.x;
  const bool in_grid = i < g.n;
  cell_xyz(g.nx, g.nxny, i, x, y, z);
  q &= 1;
  const T *__restrict__ f2 = pop<KIND>(sc.f, q);
  const T *__restrict__ h2 = pop<KIND>(sc.h, q);
  T *__restrict__ fo = pop<KIND>(sc.f, q ^ 1);
  T *__restrict__ ho = pop<KIND>(sc.h, q ^ 1);
  const float *__restrict__ phi = par(sc.phi, q);
  float *__restrict__ phi_next = par(sc.phi, q ^ 1);
  const uint8_t *__restrict__ tag = sc.tag;
  const Ctrl *__restrict__ ctrl = sc.ctrl;
  double mom[6] = {0, 0, 0, 0, 0, 0};
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = ctrl->box_fused[j];
  // Link momentum can only arise next to the sphere, inside the band box: the
  // tiles holding the box's rows store their sums (zero or not), and the body
  // update reads exactly those.
  const bool exchange = sc.has_solid != 0 && tile_in_box_rows(sc.tm, box, tile);
  const int use_bf = first && sc.use_bf;
  double cx = 0, cy = 0, cz = 0;
  if (exchange) {
    cx = ctrl->sphere.cx;
    cy = ctrl->sphere.cy;
    cz = ctrl->sphere.cz;
  }
  if (in_grid) {
    // Solid cells lie inside the band box with a margin of 2, so outside it the
    // only non-fluid cells are the walls on the grid faces.
    const bool near = in_box(box, x, y, z);
    const bool fluid = interior(g, x, y, z) && (!near || tag[i] == TAG_FLUID);
    if (!fluid) {
      if (!COLLIDE) phi_next[i] = phi[i];
    } else {
      const R phiOld = phi[i];
      const R rhoPhaseOld = R(d.rhoG) + phiOld * R(d.dRho);
      float gxo, gyo, gzo, lapo;
      phi_grad<R>(g, phi, i, phiOld, gxo, gyo, gzo, lapo);
      uint32_t nonfluid, solid;
      if (near) tag_masks(g, i, tag, nonfluid, solid);
      else {
        nonfluid = wall_mask(g, x, y, z);
        solid = 0;
      }
      float fp[Q27];
      R pSumF, mx, my, mz;
      pull27<KIND>(g, i, f2, nonfluid, solid, sc.usx, sc.usy, sc.usz, rhoPhaseOld, exchange, x, y,
                   z, cx, cy, cz, mom, fp, pSumF, mx, my, mz);
      const R rhoN = pSumF > R(0.01) ? pSumF : R(0.01);
      float hp[Q7];
      pull7<KIND>(g, i, h2, phi, nonfluid, hp);
      R phiSum = 0;
#pragma unroll
      for (int k = 0; k < Q7; k++) phiSum += (R)hp[k];
      const R p = phiSum < 0 ? R(0) : phiSum > 1 ? R(1) : phiSum;
      const R rhoPhaseNew = R(d.rhoG) + p * R(d.dRho);
      const R mu = 4 * R(d.beta) * p * (p - 1) * (2 * p - 1) - R(d.kappa) * (R)lapo;
      const R bxf = use_bf ? (R)sc.bfx[i] : R(0);
      const R byf = use_bf ? (R)sc.bfy[i] : R(0);
      const R bzf = use_bf ? (R)sc.bfz[i] : R(0);
      const R fx = bxf + mu * (R)gxo;
      const R fy = byf + mu * (R)gyo + rhoPhaseNew * R(d.gravity);
      const R fz = bzf + mu * (R)gzo;
      const R invRho = 1 / rhoPhaseNew;
      R uxn = (mx + R(0.5) * fx) * invRho;
      R uyn = (my + R(0.5) * fy) * invRho;
      R uzn = (mz + R(0.5) * fz) * invRho;
      const R u2 = uxn * uxn + uyn * uyn + uzn * uzn;
      const R UMAX = R(0.2), UMAX2 = UMAX * UMAX;
      if (u2 > UMAX2) {
        const R scl = UMAX / sqrt(u2);
        uxn *= scl;
        uyn *= scl;
        uzn *= scl;
      }
      const float rhof = (float)rhoN, phif = (float)p;
      const float uxf = (float)uxn, uyf = (float)uyn, uzf = (float)uzn;
      const bool band = near;
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
          // the reference stores them) and the gradients of phi(t+1).  The
          // post-stream populations enter as storing and reloading them would
          // give them back (exact for float32), so a cell collides the same
          // whether or not a call ends between the stream and the collide.
          if (KIND != ST_F32) {
#pragma unroll
            for (int k = 0; k < Q27; k++) fp[k] = C::rt(fp[k], w27(k));
#pragma unroll
            for (int k = 0; k < Q7; k++) hp[k] = C::rt(hp[k], w7(k) * phif);
          }
          float gx1, gy1, gz1, lap1;
          phi_grad<R>(g, phi_next, i, (R)phif, gx1, gy1, gz1, lap1);
          float hc[Q7];
          collide_phase_cell<R>(d, (R)phif, (R)uxf, (R)uyf, (R)uzf, (R)gx1, (R)gy1, (R)gz1, hp,
                                hc);
#pragma unroll
          for (int k = 0; k < Q7; k++) C::st(ho, k * g.S + i, hc[k], w7(k) * phif);
          float fc[Q27];
          collide_hydro_cell<R>(d, (R)rhof, (R)phif, (R)lap1, (R)gx1, (R)gy1, (R)gz1, R(0), R(0),
                                R(0), (R)uxf, (R)uyf, (R)uzf, fp, fc);
#pragma unroll
          for (int k = 0; k < Q27; k++) C::st(fo, k * g.S + i, fc[k], w27(k));
        }
      } else {
        phi_next[i] = phif;
      }
      if (store_stream) {
#pragma unroll
        for (int k = 0; k < Q27; k++) C::st(fo, k * g.S + i, fp[k], w27(k));
#pragma unroll
        for (int k = 0; k < Q7; k++) C::st(ho, k * g.S + i, hp[k], w7(k) * phif);
      }
    }
  }
  if (exchange) tile_momentum(mom, sc.partials, tile);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) void km_stream_collide(const Scene *__restrict__ scenes,
                                                         const int *__restrict__ off, int nscenes,
                                                         int q, int first) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  stream_collide_body<COLLIDE, KIND>(scenes[si], blockIdx.x - off[si], q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) void k1_stream_collide(const Scene sc, int q, int first) {
  stream_collide_body<COLLIDE, KIND>(sc, blockIdx.x, q, first);
}
// ---------------------------------------------------------------- storage conversion

// Float32 populations (the reference's f and h) to a 16-bit codec, and back.
// The h baseline is phi of the cell (the phi the populations belong to).
template <int KIND>
__global__ __launch_bounds__(256) void k_encode(int64_t n, int64_t S, const float *__restrict__ f,
                                                const float *__restrict__ h,
                                                const float *__restrict__ phi,
                                                typename Codec<KIND>::T *__restrict__ cf,
                                                typename Codec<KIND>::T *__restrict__ ch) {
  using C = Codec<KIND>;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float pc = phi[i];
#pragma unroll
  for (int k = 0; k < Q27; k++) C::st(cf, k * S + i, f[k * S + i], w27(k));
#pragma unroll
  for (int k = 0; k < Q7; k++) C::st(ch, k * S + i, h[k * S + i], w7(k) * pc);
}
template <int KIND>
__global__ __launch_bounds__(256) void k_decode(int64_t n, int64_t S,
                                                const typename Codec<KIND>::T *__restrict__ cf,
                                                const typename Codec<KIND>::T *__restrict__ ch,
                                                const float *__restrict__ phi,
                                                float *__restrict__ f, float *__restrict__ h) {
  using C = Codec<KIND>;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float pc = phi[i];
#pragma unroll
  for (int k = 0; k < Q27; k++) f[k * S + i] = C::ld(cf, k * S + i, w27(k));
#pragma unroll
  for (int k = 0; k < Q7; k++) h[k * S + i] = C::ld(ch, k * S + i, w7(k) * pc);
}

// Global mass check: sums over interior fluid cells of phi and of rho (the
// pressure proxy sum f), in FP64, one partial pair per 256-cell block; the host
// adds the partials in block order, so the result is deterministic.
__global__ __launch_bounds__(256) void k_mass_partials(int64_t n, Geom g,
                                                       const uint8_t *__restrict__ tag,
                                                       const float *__restrict__ phi,
                                                       const float *__restrict__ rho,
                                                       double *__restrict__ partials) {
  __shared__ double red[4][2];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double a = 0, b = 0;
  if (i < n && tag[i] == TAG_FLUID) {
    a = phi[i];
    b = rho[i];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = a;
    red[wave][1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = ((red[0][0] + red[1][0]) + red[2][0]) + red[3][0];
    partials[2 * blockIdx.x + 1] = ((red[0][1] + red[1][1]) + red[2][1]) + red[3][1];
  }
}


// Opt-in state check: per 256-cell block, whether phi, rho and u of its fluid
// cells are finite, and the largest |u| among them (FP64).
__global__ __launch_bounds__(256) void k_check_partials(int64_t n, const uint8_t *__restrict__ tag,
                                                        const float *__restrict__ phi,
                                                        const float *__restrict__ rho,
                                                        const float *__restrict__ ux,
                                                        const float *__restrict__ uy,
                                                        const float *__restrict__ uz,
                                                        double *__restrict__ partials) {
  __shared__ double red[4][2];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double bad = 0, um = 0;
  if (i < n && tag[i] == TAG_FLUID) {
    const double a = phi[i], r = rho[i], x = ux[i], y = uy[i], z = uz[i];
    const double u2 = x * x + y * y + z * z;
    if (!(a == a && r == r && u2 == u2) || fabs(a) == INFINITY || fabs(r) == INFINITY ||
        u2 == INFINITY)
      bad = 1;
    else
      um = sqrt(u2);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bad = fmax(bad, __shfl_xor(bad, off, 64));
    um = fmax(um, __shfl_xor(um, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = bad;
    red[wave][1] = um;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = fmax(fmax(red[0][0], red[1][0]), fmax(red[2][0], red[3][0]));
    partials[2 * blockIdx.x + 1] = fmax(fmax(red[0][1], red[1][1]), fmax(red[2][1], red[3][1]));
  }
}

// ---------------------------------------------------------------- body

// Body update after stream(t) of step q, one 512-thread workgroup per scene:
//   F/T += link momentum of the tiles holding this step's band box rows, in a
//          fixed order (planes dealt to waves, tiles to lanes, then lanes and
//          waves summed in order);
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
  double acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = 0;
  {
    // Plane z's tiles run from the one holding its first box cell to the one
    // holding its last, less those already counted for plane z - 1 (a tile can
    // span two planes).  Wave w takes planes box[4] + w, + 8, ...; lane l the
    // tiles lo + l, + 64, ...
    const int lane = tid & 63, wave = tid >> 6;
    for (int z = box[4] + wave; z <= box[5]; z += kBodyThreads / 64) {
      const int64_t p = (int64_t)z * tm.nxny;
      int lo = (int)((p + (int64_t)box[2] * tm.nx + box[0]) >> kTileShift);
      const int hi = (int)((p + (int64_t)box[3] * tm.nx + box[1]) >> kTileShift);
      if (z > box[4]) {
        const int hprev = (int)((p - tm.nxny + (int64_t)box[3] * tm.nx + box[1]) >> kTileShift);
        lo = lo > hprev + 1 ? lo : hprev + 1;
      }
      for (int t = lo + lane; t <= hi; t += 64) {
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] += sc.partials[(int64_t)t * 6 + k];
      }
    }
  }
  int nd = ctrl->n_dead;
  if (nd > sc.list_cap) nd = sc.list_cap;
  {
    // Dead cells: contiguous chunks of the sorted list per thread.
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
