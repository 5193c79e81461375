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

#include <type_traits>

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

// Pull offset of direction k: the source of cell i is i - dir_off(g, k).
__device__ __forceinline__ int64_t dir_off(const Geom &g, int k) {
  return (int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k];
}

__device__ __forceinline__ bool in_box(const int box[6], int x, int y, int z) {
  return x >= box[0] && x <= box[1] && y >= box[2] && y <= box[3] && z >= box[4] && z <= box[5];
}

// Gradient and Laplacian of a phi array at interior cell i (src/sim.js:186-195),
// rounded to float as the reference's gx/gy/gz/lap arrays store them.  R is the
// arithmetic type (double for the reference-exact path).
// phi of a cell and its six face neighbours.
struct Phi7 {
  float c, px, mx, py, my, pz, mz;
  // phi of the source of D3Q7 direction k (the cell itself for k = 0).
  __device__ __forceinline__ float src(int k) const {
    return k == 0 ? c : k == 1 ? mx : k == 2 ? px : k == 3 ? my : k == 4 ? py : k == 5 ? mz : pz;
  }
};
// bo4: byte offset of the cell in a float array.
__device__ __forceinline__ Phi7 load_phi7(const Geom &g, const float *__restrict__ phi,
                                          uint32_t bo4) {
  Phi7 v;
  v.c = ldb(phi, bo4);
  v.px = ldb(phi + 1, bo4);
  v.mx = ldb(phi - 1, bo4);
  v.py = ldb(phi + g.nx, bo4);
  v.my = ldb(phi - g.nx, bo4);
  v.pz = ldb(phi + g.nxny, bo4);
  v.mz = ldb(phi - g.nxny, bo4);
  return v;
}
template <typename R>
__device__ __forceinline__ void phi_grad7(const Phi7 &v, R c, float &gxo, float &gyo, float &gzo,
                                          float &lapo) {
  const R px = v.px, mx = v.mx;
  const R py = v.py, my = v.my;
  const R pz = v.pz, mz = v.mz;
  gxo = (float)(R(0.5) * (px - mx));
  gyo = (float)(R(0.5) * (py - my));
  gzo = (float)(R(0.5) * (pz - mz));
  lapo = (float)(px + mx + py + my + pz + mz - 6 * c);
}
template <typename R>
__device__ __forceinline__ void phi_grad(const Geom &g, const float *__restrict__ phi, int64_t i,
                                         R c, float &gxo, float &gyo, float &gzo, float &lapo) {
  phi_grad7<R>(load_phi7(g, phi, (uint32_t)i * 4u), c, gxo, gyo, gzo, lapo);
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

// ---------------------------------------------------------------- 16-bit codecs: deviation form
//
// With 16-bit storage the step computes in FP32 on the stored deviations
// themselves: d_k = f_k - w_k, and h_k - w7_k phi with phi that of the cell the
// population is stored at.  A pulled f deviation is exactly the stored value,
// so decoding is one conversion and a post-stream population goes back to
// storage without loss; the momentum needs no reference level (w_k e_k sums to
// zero over each opposite pair); the equilibrium enters as its deviation
// w_k (rho - 1 + rho_phase Q_k), and products and sums are fused (fmaf) where
// that saves an instruction.  The results are not the float32 form's bit for
// bit (that form is FP64 on full values) but are deterministic, and the fused
// step and a call split between stream and collide store and reload the same
// 16-bit values.  D3Q27 directions come in opposite pairs (2j - 1, 2j), so e.v
// is formed once per pair.

// Hydro collide in deviation form (collideHydro, src/sim.js:430-459): fin(k)
// gives d_k = f_k - w_k of the cell and out(k, v) receives the post-collision
// deviation of direction k; rhof is rho as stored (sum f).
//
// The BGK step with Guo forcing, d_k + (1/tau)(w_k (rho - 1 + rho_phase Q_k)
// - d_k) + A w_k ((e_k - u).F / cs^2 + (e_k.u)(e_k.F) / cs^4), with
// Q_k = 3 e.u + 4.5 (e.u)^2 - 1.5 u^2, is rearranged as
//   (1 - 1/tau) d_k + w_k (c0 + c1 e.u + c2 (e.u)^2 + A e.F + 3 A (e.u)(e.F))
// with per-cell c0 = (rho_phase (-1.5 u^2) + rho - 1) / tau - A u.F,
// c1 = 3 rho_phase / tau, c2 = 4.5 rho_phase / tau.  For an opposite pair
// (e, -e) the bracket is sym +- anti with sym = c0 + e.u (c2 e.u + 3 A e.F) and
// anti = c1 e.u + A e.F, and the weight w_k is folded into the constants of
// its class (faces, edges, corners): about 10 operations per pair instead of 18.
template <int K, typename FIn, typename FOut>
__device__ __forceinline__ void hydro_dev_pair(float c0, float c1, float c2, float om, float ux,
                                               float uy, float uz, float ax, float ay, float az,
                                               FIn &fin, FOut &out) {
  // K: the first direction of a pair; K + 1 is its opposite.  c0..c2 and
  // (ax, ay, az) = A F carry the pair's weight.
  constexpr int ex = EX27[K], ey = EY27[K], ez = EZ27[K];
  const float eu = edot<ex, ey, ez>(ux, uy, uz);
  const float aF = edot<ex, ey, ez>(ax, ay, az);
  const float sym = fmaf(eu, fmaf(c2, eu, 3.0f * aF), c0);
  const float anti = fmaf(c1, eu, aF);
  out(K, fmaf(om, fin(K), sym + anti));
  out(K + 1, fmaf(om, fin(K + 1), sym - anti));
}
// Pairs K, K + 2, ... below KEND, all of one weight.
template <int K, int KEND, typename FIn, typename FOut>
__device__ __forceinline__ void hydro_dev_pairs(float c0, float c1, float c2, float om, float ux,
                                                float uy, float uz, float ax, float ay, float az,
                                                FIn &fin, FOut &out) {
  if constexpr (K < KEND) {
    static_assert(w27(K) == w27(KEND - 1), "one weight class");
    hydro_dev_pair<K>(c0, c1, c2, om, ux, uy, uz, ax, ay, az, fin, out);
    hydro_dev_pairs<K + 2, KEND>(c0, c1, c2, om, ux, uy, uz, ax, ay, az, fin, out);
  }
}
// The pairs of directions K .. KEND - 1 with the class weight folded in.
template <int K, int KEND, typename FIn, typename FOut>
__device__ __forceinline__ void hydro_dev_class(float c0, float c1, float c2, float om, float ux,
                                                float uy, float uz, float afx, float afy,
                                                float afz, FIn &fin, FOut &out) {
  constexpr float w = w27(K);
  hydro_dev_pairs<K, KEND>(w * c0, w * c1, w * c2, om, ux, uy, uz, w * afx, w * afy, w * afz,
                           fin, out);
}
template <typename FIn, typename FOut>
__device__ __forceinline__ void collide_hydro_dev(const DerivedF &d, float rhof, float p, float lap,
                                                  float ggx, float ggy, float ggz, float bx,
                                                  float by, float bz, float ux, float uy, float uz,
                                                  FIn &&fin, FOut &&out) {
  const float rhoPhase = d.rhoG + p * d.dRho;
  const float mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * lap;
  const float fx = bx + mu * ggx;
  const float fy = by + mu * ggy + rhoPhase * d.gravity;
  const float fz = bz + mu * ggz;
  const float nu2h = -1.5f * (ux * ux + uy * uy + uz * uz);
  const float uF = ux * fx + uy * fy + uz * fz;
  const float it = d.invTau;
  const float c0 = fmaf(it, fmaf(rhoPhase, nu2h, rhof - 1.0f), -d.A * uF);
  const float itr = it * rhoPhase;
  const float c1 = float(INV_CS2_27) * itr, c2 = float(INV_2CS4_27) * itr;
  const float om = 1.0f - it;
  const float afx = d.A * fx, afy = d.A * fy, afz = d.A * fz;
  // k = 0: e = 0, so the bracket is c0.
  out(0, fmaf(om, fin(0), w27(0) * c0));
  hydro_dev_class<1, 7>(c0, c1, c2, om, ux, uy, uz, afx, afy, afz, fin, out);
  hydro_dev_class<7, 19>(c0, c1, c2, om, ux, uy, uz, afx, afy, afz, fin, out);
  hydro_dev_class<19, 27>(c0, c1, c2, om, ux, uy, uz, afx, afy, afz, fin, out);
}

// Phase collide in deviation form (collidePhase, src/sim.js:341-360): hin and
// hout are h_k - w7_k p, p the phi of the cell (the collide keeps it), so
// h_eq - w7_k p = w7_k p e_k.u / cs^2.
__device__ __forceinline__ void collide_phase_dev(const DerivedF &d, float p, float ux, float uy,
                                                  float uz, float ggx, float ggy, float ggz,
                                                  const float hin[Q7], float hout[Q7]) {
  const float gmag = sqrtf(ggx * ggx + ggy * ggy + ggz * ggz) + 1e-12f;
  const float sf = d.sharpC * p * (1 - p) * d.sharpFac / gmag;
  const float S[3] = {sf * ggx, sf * ggy, sf * ggz};
  const float u[3] = {ux, uy, uz};
  const float it = d.invTauPhi;
  const float a = w7(1) * p * float(INV_CS2_7);
  hout[0] = fmaf(it, -hin[0], hin[0]);
#pragma unroll
  for (int k = 1; k < Q7; k++) {
    const int ax = (k - 1) >> 1;
    const float sg = (k & 1) ? 1.0f : -1.0f;
    const float diff = fmaf(a, sg * u[ax], -hin[k]);
    hout[k] = fmaf(w7(k) * sg, S[ax], fmaf(it, diff, hin[k]));
  }
}

// phi(t+1) of a fluid cell from its pulled h deviations e and their baselines'
// phi (the source's, or the cell's own for a bounced population): the full
// populations, summed in direction order, and their sum clamped to [0, 1].
// The phase stream and the stream kernel both use it, so both write the same
// phi(t+1).
__device__ __forceinline__ float phase_sum_dev(const float e[Q7], const float phib[Q7],
                                               float hfull[Q7]) {
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < Q7; k++) {
    hfull[k] = fmaf(w7(k), phib[k], e[k]);
    sum += hfull[k];
  }
  return sum < 0.0f ? 0.0f : sum > 1.0f ? 1.0f : sum;
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
  const uint32_t bo = (uint32_t)i;
#pragma unroll
  for (int k = 1; k < Q27; k++) {
    const uint8_t st = ldb(tag - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]), bo);
    nonfluid |= (uint32_t)(st != TAG_FLUID) << k;
    solid |= (uint32_t)(st == TAG_SOLID) << k;
  }
}

// Halfway bounce-back and momentum exchange of the 27 pulled populations
// (src/sim.js:509-543): raw[k] is the population pulled from the source cell,
// or the cell's own opposite population where the source is not fluid; the
// moving-wall term and the momentum exchange run only for solid links.  fk
// stays the unrounded value the reference sums and exchanges; fp receives it
// rounded to float.  The momentum exchange accumulates in FP64 whatever the
// arithmetic type R.
template <typename R>
__device__ __forceinline__ void links27(const Geom &g, int64_t i, const float raw[Q27],
                                        uint32_t solid, const float *__restrict__ usx,
                                        const float *__restrict__ usy,
                                        const float *__restrict__ usz, R rhoPhaseOld,
                                        bool exchange, int x, int y, int z, double cx, double cy,
                                        double cz, double mom[6], float fp[Q27], R &pSumF, R &mx,
                                        R &my, R &mz) {
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

// Element pulled for direction k of a cell at byte offset bo = i * sizeof(T):
// the source plane's element i - off, or at a non-fluid source the cell's own
// element of the opposite direction.  A direction and its opposite hold
// adjacent planes (k odd: k, k + 1), so both candidates lie within two planes
// of the uniform base at the lower of them, and the choice is a 32-bit byte
// offset made per lane without branching: the address needs no 64-bit
// arithmetic and all pulls of a cell are in flight together.  Needs
// 2 S sizeof(T) < 2^32 (float32 scenes: S < 2^29), and for the source i >= off,
// which holds for fluid cells (their sources lie in the grid).
template <typename T>
__device__ __forceinline__ T pull_elem(const T *__restrict__ p2, int64_t S, int k, int off,
                                       uint32_t bo, bool bounce) {
  if (k == 0) return ldb(p2, bo);
  const bool lower = (k & 1) != 0;  // k is the lower plane of its pair
  const T *base = p2 + (int64_t)(lower ? k : k - 1) * S;
  const uint32_t plane = (uint32_t)S * (uint32_t)sizeof(T);
  const uint32_t src = bo - (uint32_t)off * (uint32_t)sizeof(T) + (lower ? 0u : plane);
  const uint32_t own = bo + (lower ? plane : 0u);
  return ldb(base, bounce ? own : src);
}

// Float32 populations of the cells of one wave: a wave-uniform element base
// and each lane's 32-bit byte offset from it to its own cell's direction-0
// element, so that every load and store adds one per-lane offset to a uniform
// base and needs no 64-bit address arithmetic per lane.  Pulls use one base per
// buffer, D elements below the wave's own (D: the largest distance of a source
// element below a cell's own, pull_span), so every direction's two candidates
// are 32-bit byte offsets above it.
struct RowAddr {
  int64_t base;
  uint32_t relb;
};
// i, r: the lane's cell and its row (z ny + y).  The base is the element of the
// wave's first active lane (the lowest cell, in the lowest row, among the
// active lanes), read from that lane, so nothing is divided.
__device__ __forceinline__ RowAddr row_addr(const Geom &g, int64_t i, int r) {
  const int ii = (int)i;
  const int i0 = __builtin_amdgcn_readfirstlane(ii);
  const int r0 = __builtin_amdgcn_readfirstlane(r);
  const int pa = (int)g.pa;
  RowAddr a;
  a.base = (int64_t)i0 + (int64_t)r0 * pa;
  a.relb = (uint32_t)((ii - i0) + (r - r0) * pa) * 4u;
  return a;
}
// Direction k (opposite kopp, source (x - ex, y - ey, z - ez)) of a cell from
// lo = the population pointer plus the wave base minus D elements: the source's
// element, or at a non-fluid source the cell's own element of the opposite
// direction; the per-lane choice is between two uniform 32-bit offsets.  Needs
// f32_fused_fits (checked on the host) and the source inside the grid, which
// holds for fluid cells.
__device__ __forceinline__ float pull_row(const float *__restrict__ lo, uint32_t relb,
                                          const Geom &g, int k, int kopp, int ex, int ey, int ez,
                                          bool bounce) {
  const uint32_t pb = (uint32_t)g.pb, D = (uint32_t)pull_span(g);
  const uint32_t delta = (uint32_t)((ez * g.ny + ey) * (g.nx + (int)g.pa) + ex);
  const uint32_t ds = ((uint32_t)k * pb - delta + D) * 4u;
  const uint32_t dop = ((uint32_t)kopp * pb + D) * 4u;
  return ldb(lo, relb + (bounce ? dop : ds));
}
// The pull base of a population pointer for a wave.
__device__ __forceinline__ const float *pull_base(const float *p, const RowAddr &a, const Geom &g) {
  return p + (a.base - pull_span(g));
}

// Pull stream of the 27 populations with halfway bounce-back, given the source
// masks.  Every load address is chosen without branching (own opposite
// population at a non-fluid source), so the 27 loads are in flight together.
template <int KIND>
__device__ __forceinline__ void pull27(const Geom &g, int64_t i, const RowAddr &ra,
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
  static_assert(KIND == ST_F32, "16-bit codecs pull in deviation form");
  using C = Codec<KIND>;
  float raw[Q27];
  const float *lo = pull_base(f2, ra, g);
#pragma unroll
  for (int k = 0; k < Q27; k++)
    raw[k] = pull_row(lo, ra.relb, g, k, opp27(k), EX27[k], EY27[k], EZ27[k], (nonfluid >> k) & 1u);
  links27<typename C::R>(g, i, raw, solid, usx, usy, usz, rhoPhaseOld, exchange, x, y, z, cx, cy,
                         cz, mom, fp, pSumF, mx, my, mz);
}

// links27 in deviation form (16-bit codecs): get(k) gives the pulled
// deviation (the cell's own opposite one at a non-fluid source; w_k is the same
// for a direction and its opposite).  A solid link adds the moving-wall term
// and hands the corrected deviation to fix(k, d): the collide's input is then
// that deviation as storing and reloading it gives it back (the others are
// exact already).  sumD and the momentum sum the unrounded deviations.
template <int KIND, typename Get, typename Fix>
__device__ __forceinline__ void links27_dev(const Geom &g, int64_t i, Get &&get, Fix &&fix,
                                            uint32_t solid, const float *__restrict__ usx,
                                            const float *__restrict__ usy,
                                            const float *__restrict__ usz, float rhoPhaseOld,
                                            bool exchange, int x, int y, int z, double cx,
                                            double cy, double cz, double mom[6], float &sumD,
                                            float &mx, float &my, float &mz) {
  sumD = 0;
  mx = my = mz = 0;
  // Impulse of the cell's solid links, -sum m_k e_k (applyImpulseAtCell at
  // x - e_k with -m_k e_k, m_k the sum of the two full populations).  Their
  // torque about the sphere centre is sum (r0 - e_k) x (-m_k e_k) = r0 x P
  // exactly (e_k x e_k = 0), r0 the cell centre's offset: one cross product
  // per cell.
  double Px = 0, Py = 0, Pz = 0;
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    float dk = get(k);
    if ((solid >> k) & 1u) {
      const int ex = EX27[k], ey = EY27[k], ez = EZ27[k];
      const int64_t s = i - ((int64_t)ez * g.nxny + (int64_t)ey * g.nx + ex);
      const float dopp = dk;
      const float eu = (float)ex * usx[s] + (float)ey * usy[s] + (float)ez * usz[s];
      dk = dopp + 2 * w27(k) * rhoPhaseOld * eu * float(INV_CS2_27);
      fix(k, dk);
      if (exchange) {
        const double m = 2.0 * (double)w27(k) + ((double)dopp + (double)dk);
        if (EX27[k] > 0) Px -= m;
        if (EX27[k] < 0) Px += m;
        if (EY27[k] > 0) Py -= m;
        if (EY27[k] < 0) Py += m;
        if (EZ27[k] > 0) Pz -= m;
        if (EZ27[k] < 0) Pz += m;
      }
    }
    sumD += dk;
    if (EX27[k] > 0) mx += dk;
    if (EX27[k] < 0) mx -= dk;
    if (EY27[k] > 0) my += dk;
    if (EY27[k] < 0) my -= dk;
    if (EZ27[k] > 0) mz += dk;
    if (EZ27[k] < 0) mz -= dk;
  }
  if (exchange && solid) {
    const double rx = x + 0.5 - cx, ry = y + 0.5 - cy, rz = z + 0.5 - cz;
    mom[0] += Px;
    mom[1] += Py;
    mom[2] += Pz;
    mom[3] += ry * Pz - rz * Py;
    mom[4] += rz * Px - rx * Pz;
    mom[5] += rx * Py - ry * Px;
  }
}

// Pull of the 7 float32 phase populations (src/sim.js:376-397): from the
// source cell, or the own opposite population when the source is not fluid
// (D3Q7 directions 1..6 are D3Q27 directions 1..6, so the D3Q27 source mask
// serves).
template <int KIND>
__device__ __forceinline__ void pull7(const Geom &g, const RowAddr &ra,
                                      const typename Codec<KIND>::T *__restrict__ h2,
                                      uint32_t nonfluid, float hp[Q7]) {
  static_assert(KIND == ST_F32, "16-bit codecs pull in deviation form");
  const float *lo = pull_base(h2, ra, g);
#pragma unroll
  for (int k = 0; k < Q7; k++)
    hp[k] = pull_row(lo, ra.relb, g, k, opp7(k), EX7[k], EY7[k], EZ7[k], (nonfluid >> k) & 1u);
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
  DerivedF df;  // d rounded to float (16-bit codecs)
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
  int band_blocks;          // blocks of 256 cells covering the largest band box
  int band_tile_cols;       // band-row tiles per plane a band-tile launch covers
  int band_tile_planes;     // planes it covers (its blocks: cols * planes)
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

// Cell of block lb of a launch over the whole grid (blocks of 64x4 cells) or
// over the band box (blocks of 256 box cells); false when outside.
template <bool BAND>
__device__ __forceinline__ bool block_cell(const Scene &sc, int lb, int &x, int &y, int &z) {
  if (BAND) {
    // The box's cells numbered x fastest, 256 per block, so narrow boxes
    // leave no lanes idle.
    const int *box = sc.ctrl->box_band;
    const int ex = box[1] - box[0] + 1, ey = box[3] - box[2] + 1, ez = box[5] - box[4] + 1;
    if (ex <= 0 || ey <= 0 || ez <= 0) return false;
    const int t = lb * 256 + (int)threadIdx.y * 64 + (int)threadIdx.x;
    if (t >= ex * ey * ez) return false;
    int r;
    z = box[4] + small_divmod(t, ex * ey, r);
    y = box[2] + small_divmod(r, ex, x);
    x += box[0];
    return true;
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
  const uint32_t bo = (uint32_t)i * (uint32_t)sizeof(T);
  if constexpr (KIND != ST_F32) {
    // Deviation form: h is stored against w7_k phi of this cell, f against w_k.
    const uint32_t bo4 = (uint32_t)i * 4u;
    const float uxf = ldb(sc.ux, bo4), uyf = ldb(sc.uy, bo4), uzf = ldb(sc.uz, bo4);
    float hin[Q7], hc[Q7];
#pragma unroll
    for (int k = 0; k < Q7; k++) hin[k] = C::ldd(hi + (int64_t)k * g.S, bo);
    collide_phase_dev(sc.df, phc, uxf, uyf, uzf, gx1, gy1, gz1, hin, hc);
#pragma unroll
    for (int k = 0; k < Q7; k++) C::std_(ho + (int64_t)k * g.S, bo, hc[k]);
    float fin[Q27];
#pragma unroll
    for (int k = 0; k < Q27; k++) fin[k] = C::ldd(fi + (int64_t)k * g.S, bo);
    const bool bf = !BAND && sc.use_bf;
    const float bx = bf ? ldb(sc.bfx, bo4) : 0.0f;
    const float by = bf ? ldb(sc.bfy, bo4) : 0.0f;
    const float bz = bf ? ldb(sc.bfz, bo4) : 0.0f;
    collide_hydro_dev(sc.df, ldb(sc.rho, bo4), phc, lap1, gx1, gy1, gz1, bx, by, bz, uxf, uyf, uzf,
                      [&](int k) { return fin[k]; },
                      [&](int k, float v) { C::std_(fo + (int64_t)k * g.S, bo, v); });
    return;
  }
  // Float32: row blocks.
  const int64_t pe = pop_base(g, i);
  float hin[Q7], hc[Q7];
#pragma unroll
  for (int k = 0; k < Q7; k++) hin[k] = C::ldo(hi + pe + k * g.pb, 0u, w7(k) * phc);
  collide_phase_cell<R>(sc.d, pc, uxe, uye, uze, (R)gx1, (R)gy1, (R)gz1, hin, hc);
#pragma unroll
  for (int k = 0; k < Q7; k++) C::sto(ho + pe + k * g.pb, 0u, hc[k], w7(k) * phc);
  float fin[Q27], fc[Q27];
#pragma unroll
  for (int k = 0; k < Q27; k++) fin[k] = C::ldo(fi + pe + k * g.pb, 0u, w27(k));
  const bool bf = !BAND && sc.use_bf;
  const R bx = bf ? (R)sc.bfx[i] : R(0);
  const R by = bf ? (R)sc.bfy[i] : R(0);
  const R bz = bf ? (R)sc.bfz[i] : R(0);
  collide_hydro_cell<R>(sc.d, (R)sc.rho[i], pc, (R)lap1, (R)gx1, (R)gy1, (R)gz1, bx, by, bz, uxe,
                        uye, uze, fin, fc);
#pragma unroll
  for (int k = 0; k < Q27; k++) C::sto(fo + pe + k * g.pb, 0u, fc[k], w27(k));
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
  const uint32_t bo4 = (uint32_t)i * 4u;
  if (!interior(g, x, y, z) || (near && ldb(tag, (uint32_t)i) != TAG_FLUID)) {
    stb(phi_next, bo4, ldb(phi, bo4));
    return;
  }
  uint32_t nonfluid = 0;
  if (near) {
#pragma unroll
    for (int k = 1; k < Q7; k++) {
      const int64_t off = (int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k];
      nonfluid |= (uint32_t)(ldb(tag - off, (uint32_t)i) != TAG_FLUID) << k;
    }
  } else {
    nonfluid = wall_mask(g, x, y, z) & 0x7eu;  // D3Q7 directions 1..6
  }
  if constexpr (KIND != ST_F32) {
    // The stream kernel's phi(t+1) of the cell (phase_sum_dev), from the same
    // pulled deviations and baselines.
    // Each address chosen without branching (own opposite population at a
    // non-fluid source): one load per direction, all in flight together, also
    // in waves that hold wall cells (most waves of a narrow grid).
    using C = Codec<KIND>;
    const Phi7 v = load_phi7(g, phi, bo4);
    float e[Q7], phib[Q7], hfull[Q7];
    const uint16_t *h16 = reinterpret_cast<const uint16_t *>(h2);
#pragma unroll
    for (int k = 0; k < Q7; k++) {
      const bool bounce = (nonfluid >> k) & 1u;
      e[k] = C::dv(pull_elem(h16, g.S, k, (int)dir_off(g, k), (uint32_t)i * 2u, bounce));
      phib[k] = bounce ? v.c : v.src(k);
    }
    stb(phi_next, bo4, phase_sum_dev(e, phib, hfull));
  } else {
    float hv[Q7];
    pull7<KIND>(g, row_addr(g, i, z * g.ny + y), h2, nonfluid, hv);
    R phiSum = 0;
#pragma unroll
    for (int k = 0; k < Q7; k++) phiSum += (R)hv[k];
    const R pp = phiSum < 0 ? R(0) : phiSum > 1 ? R(1) : phiSum;
    stb(phi_next, bo4, (float)pp);
  }
}
// Tile launches over a scene set: tiles[b] = (scene, tile of the scene) of block b.
template <int KIND>
__global__ __launch_bounds__(256) void km_phase_stream(const Scene *__restrict__ scenes,
                                                       const int2 *__restrict__ tiles, int q) {
  const int2 t = tiles[blockIdx.x];
  phase_stream_body<KIND>(scenes[t.x], t.y, q);
}
template <int KIND>
__global__ __launch_bounds__(256) void k1_phase_stream(const Scene sc, int q) {
  phase_stream_body<KIND>(sc, blockIdx.x, q);
}

// ---------------------------------------------------------------- stream + collide

// Tile of block b of a launch over the band-row tiles (the tiles that hold a
// cell of the band box's rows, as tile_in_box_rows finds them and the body
// update reads their momentum): block b takes plane box[4] + b / M and, in it,
// the (b % M)-th of the plane's band-row tiles not already counted for the
// plane before (a tile can span two planes).  M = band_tile_cols; the host
// sizes the launch (band_tile_planes planes of M blocks) for the largest box.
// A plane with more band-row tiles than M, or a box with more planes than the
// launch covers, sets kErrBandTiles.
__device__ __forceinline__ bool band_tile(const Scene &sc, const int box[6], int b, int &tile) {
  const int M = sc.band_tile_cols;
  if (box[5] < box[4]) return false;
  if (b == 0 && box[5] - box[4] + 1 > sc.band_tile_planes) atomicOr(&sc.ctrl->err, kErrBandTiles);
  const int z = box[4] + b / M;
  if (z > box[5]) return false;
  const TileMap &tm = sc.tm;
  const int64_t p = (int64_t)z * tm.nxny;
  int lo = (int)((p + (int64_t)box[2] * tm.nx + box[0]) >> kTileShift);
  const int hi = (int)((p + (int64_t)box[3] * tm.nx + box[1]) >> kTileShift);
  if (z > box[4]) {
    const int hprev = (int)((p - tm.nxny + (int64_t)box[3] * tm.nx + box[1]) >> kTileShift);
    lo = lo > hprev + 1 ? lo : hprev + 1;
  }
  const int c = b % M;
  if (c == M - 1 && lo + c < hi) atomicOr(&sc.ctrl->err, kErrBandTiles);
  tile = lo + c;
  return tile <= hi;
}

// stream(t) + macro(t) for every fluid cell of step q; with COLLIDE, also
// collide(t+1) for fluid cells outside the band.  Without COLLIDE (the last
// step of a call) the post-stream populations are stored for every cell, which
// leaves the state exactly as the reference's step does, and phi(t+1) is
// written here.  first: the first step of the call (the armed external force
// applies to its macroscopic update).
// Macro(t) and either collide(t+1) or the post-stream populations of fluid cell
// i, from its pulled populations (fp after bounce-back, with their sums pSumF
// and momentum mx, my, mz; hp) and the gradients of phi(t) at i.  band: the cell
// lies in the band box (its collide(t+1) runs after the band update).  The 27 +
// 7 populations to store are handed to stf(k, value, baseline) and
// sth(k, value, baseline).  phi1: the stencil of phi(t+1) around the cell when
// the caller holds it (the row march), else it is loaded from phi_next.
template <bool COLLIDE, int KIND, typename StF, typename StH>
__device__ __forceinline__ void finish_cell(const Scene &sc, int64_t i, bool band,
                                            const float *__restrict__ phi_next, float gxo,
                                            float gyo, float gzo, float lapo, float fp[Q27],
                                            float hp[Q7], typename Codec<KIND>::R pSumF,
                                            typename Codec<KIND>::R mx,
                                            typename Codec<KIND>::R my,
                                            typename Codec<KIND>::R mz, int use_bf,
                                            float *__restrict__ phi_out, StF &&stf, StH &&sth,
                                            const Phi7 *phi1 = nullptr) {
  using C = Codec<KIND>;
  using R = typename C::R;
  const Geom &g = sc.g;
  const Derived &d = sc.d;
  const R rhoN = pSumF > R(0.01) ? pSumF : R(0.01);
  R phiSum = 0;
#pragma unroll
  for (int k = 0; k < Q7; k++) phiSum += (R)hp[k];
  const R p = phiSum < 0 ? R(0) : phiSum > 1 ? R(1) : phiSum;
  const R rhoPhaseNew = R(d.rhoG) + p * R(d.dRho);
  const R mu = 4 * R(d.beta) * p * (p - 1) * (2 * p - 1) - R(d.kappa) * (R)lapo;
  const uint32_t bo4 = (uint32_t)i * 4u;
  const R bxf = use_bf ? (R)ldb(sc.bfx, bo4) : R(0);
  const R byf = use_bf ? (R)ldb(sc.bfy, bo4) : R(0);
  const R bzf = use_bf ? (R)ldb(sc.bfz, bo4) : R(0);
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
  // rho and u are read before the end of the call only inside the band (by
  // the fresh-cell pass, the dead-cell impulses and the band collide); the
  // last step of a call stores them everywhere.
  if (!COLLIDE || band) {
    stb(sc.rho, bo4, rhof);
    stb(sc.ux, bo4, uxf);
    stb(sc.uy, bo4, uyf);
    stb(sc.uz, bo4, uzf);
  }
  if (COLLIDE && !band) {
    // collide(t+1) with the new macroscopic values (rounded to float as the
    // reference stores them) and the gradients of phi(t+1).  The post-stream
    // populations enter as storing and reloading them would give them back
    // (exact for float32), so a cell collides the same whether or not a call
    // ends between the stream and the collide.
    if (KIND != ST_F32) {
#pragma unroll
      for (int k = 0; k < Q27; k++) fp[k] = C::rt(fp[k], w27(k));
#pragma unroll
      for (int k = 0; k < Q7; k++) hp[k] = C::rt(hp[k], w7(k) * phif);
    }
    float gx1, gy1, gz1, lap1;
    if (phi1) phi_grad7<R>(*phi1, (R)phif, gx1, gy1, gz1, lap1);
    else phi_grad<R>(g, phi_next, i, (R)phif, gx1, gy1, gz1, lap1);
    float hc[Q7];
    collide_phase_cell<R>(d, (R)phif, (R)uxf, (R)uyf, (R)uzf, (R)gx1, (R)gy1, (R)gz1, hp, hc);
#pragma unroll
    for (int k = 0; k < Q7; k++) sth(k, hc[k], w7(k) * phif);
    float fc[Q27];
    collide_hydro_cell<R>(d, (R)rhof, (R)phif, (R)lap1, (R)gx1, (R)gy1, (R)gz1, R(0), R(0), R(0),
                          (R)uxf, (R)uyf, (R)uzf, fp, fc);
#pragma unroll
    for (int k = 0; k < Q27; k++) stf(k, fc[k], w27(k));
  } else {
    if (!COLLIDE) stb(phi_out, bo4, phif);
#pragma unroll
    for (int k = 0; k < Q27; k++) stf(k, fp[k], w27(k));
#pragma unroll
    for (int k = 0; k < Q7; k++) sth(k, hp[k], w7(k) * phif);
  }
}

// finish_cell in deviation form (16-bit codecs): fin(k) gives the collide
// input f_k - w_k, e holds the pulled h deviations with their baselines' phi
// phib, sumD and mx, my, mz the deviation sums.  stf(k, dev) and sth(k, dev)
// receive the deviations to store (h against w7_k phi(t+1) of the cell); stf
// for direction k follows fin(k).
template <bool COLLIDE, int KIND, typename FIn, typename StF, typename StH>
__device__ __forceinline__ void finish_cell_dev(const Scene &sc, int64_t i, bool band,
                                                const float *__restrict__ phi_next, float gxo,
                                                float gyo, float gzo, float lapo, FIn &&fin,
                                                const float e[Q7], const float phib[Q7], float sumD,
                                                float mx, float my, float mz, int use_bf,
                                                float *__restrict__ phi_out, StF &&stf,
                                                StH &&sth) {
  using C = Codec<KIND>;
  const Geom &g = sc.g;
  const DerivedF &d = sc.df;
  const float pSum = 1.0f + sumD;
  const float rhoN = pSum > 0.01f ? pSum : 0.01f;
  float h[Q7];
  const float p = phase_sum_dev(e, phib, h);
  const float rhoPhaseNew = d.rhoG + p * d.dRho;
  const float mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * lapo;
  const uint32_t bo4 = (uint32_t)i * 4u;
  const float bxf = use_bf ? ldb(sc.bfx, bo4) : 0.0f;
  const float byf = use_bf ? ldb(sc.bfy, bo4) : 0.0f;
  const float bzf = use_bf ? ldb(sc.bfz, bo4) : 0.0f;
  const float fx = bxf + mu * gxo;
  const float fy = byf + mu * gyo + rhoPhaseNew * d.gravity;
  const float fz = bzf + mu * gzo;
  const float invRho = 1 / rhoPhaseNew;
  float uxn = (mx + 0.5f * fx) * invRho;
  float uyn = (my + 0.5f * fy) * invRho;
  float uzn = (mz + 0.5f * fz) * invRho;
  const float u2 = uxn * uxn + uyn * uyn + uzn * uzn;
  constexpr float UMAX = 0.2f;
  if (u2 > UMAX * UMAX) {
    const float scl = UMAX / sqrtf(u2);
    uxn *= scl;
    uyn *= scl;
    uzn *= scl;
  }
  const float rhof = rhoN;
  if (!COLLIDE || band) {
    stb(sc.rho, bo4, rhof);
    stb(sc.ux, bo4, uxn);
    stb(sc.uy, bo4, uyn);
    stb(sc.uz, bo4, uzn);
  }
  // h against the cell's new phi, the baseline it is stored with.
  float hd[Q7];
#pragma unroll
  for (int k = 0; k < Q7; k++) hd[k] = fmaf(-w7(k), p, h[k]);
  if (COLLIDE && !band) {
#pragma unroll
    for (int k = 0; k < Q7; k++) hd[k] = C::rtv(hd[k]);
    float gx1, gy1, gz1, lap1;
    phi_grad<float>(g, phi_next, i, p, gx1, gy1, gz1, lap1);
    float hc[Q7];
    collide_phase_dev(d, p, uxn, uyn, uzn, gx1, gy1, gz1, hd, hc);
#pragma unroll
    for (int k = 0; k < Q7; k++) sth(k, hc[k]);
    collide_hydro_dev(d, rhof, p, lap1, gx1, gy1, gz1, 0.0f, 0.0f, 0.0f, uxn, uyn, uzn, fin, stf);
  } else {
    if (!COLLIDE) stb(phi_out, bo4, p);
#pragma unroll
    for (int k = 0; k < Q27; k++) stf(k, fin(k));
#pragma unroll
    for (int k = 0; k < Q7; k++) sth(k, hd[k]);
  }
}

// stream(t) + macro(t) for every fluid cell, then collide(t+1) outside the
// band; one cell per lane, tiles of 256 cells.  COLLIDE false: the stream-only
// step that ends a call (phi(t+1) written here, rho/u stored everywhere).
// MODE: kTilesAll covers every tile; kTilesBulk every tile but the band-row
// tiles, whose cells all lie outside the band box (no tag is read, no link
// momentum arises: the kernel carries neither path); kTilesBand only the
// band-row tiles (band_tile).  Bulk and band launches together cover every
// tile once.  kTilesBox: the band-row tiles' cells inside the band box only
// (the row march updates every other cell); lanes outside it add nothing to
// the tile's momentum, as in kTilesBand, so both give the same partials.
constexpr int kTilesAll = 0, kTilesBulk = 1, kTilesBand = 2, kTilesBox = 3;
template <bool COLLIDE, int KIND, int MODE = kTilesAll>
__device__ __forceinline__ void stream_collide_body(const Scene &sc, int lb, int q, int first) {
  using C = Codec<KIND>;
  using R = typename C::R;
  using T = typename C::T;
  const Geom &g = sc.g;
  const Derived &d = sc.d;
  const Ctrl *__restrict__ ctrl = sc.ctrl;
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = ctrl->box_fused[j];
  int tile, x, y, z;
  if constexpr (MODE == kTilesBand || MODE == kTilesBox) {
    if (!band_tile(sc, box, lb, tile)) return;
  } else {
    if (!map_tile(sc.tm.T, sc.tm.T8, lb, tile)) return;
  }
  // Link momentum can only arise next to the sphere, inside the band box: the
  // tiles holding the box's rows store their sums (zero or not), and the body
  // update reads exactly those.
  const bool box_rows = MODE == kTilesBand || MODE == kTilesBox ||
                        (MODE == kTilesAll && tile_in_box_rows(sc.tm, box, tile));
  if constexpr (MODE == kTilesBulk) {
    if (tile_in_box_rows(sc.tm, box, tile)) return;
  }
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
  double mom[6] = {0, 0, 0, 0, 0, 0};
  const bool exchange = sc.has_solid != 0 && box_rows;
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
    const bool near = MODE != kTilesBulk && in_box(box, x, y, z);
    const uint32_t bo4 = (uint32_t)i * 4u;
    const bool fluid = (MODE != kTilesBox || near) && interior(g, x, y, z) &&
                       (!near || ldb(tag, (uint32_t)i) == TAG_FLUID);
    if (!fluid) {
      if (!COLLIDE) stb(phi_next, bo4, ldb(phi, bo4));
    } else if constexpr (KIND != ST_F32) {
      const Phi7 v = load_phi7(g, phi, bo4);
      const float phiOld = v.c;
      const float rhoPhaseOld = sc.df.rhoG + phiOld * sc.df.dRho;
      float gxo, gyo, gzo, lapo;
      phi_grad7<float>(v, phiOld, gxo, gyo, gzo, lapo);
      uint32_t nonfluid, solid;
      if (near) tag_masks(g, i, tag, nonfluid, solid);
      else {
        nonfluid = wall_mask(g, x, y, z);
        solid = 0;
      }
      const uint32_t bo = (uint32_t)i * (uint32_t)sizeof(T);
      float raw[Q27];
#pragma unroll
      for (int k = 0; k < Q27; k++)
        raw[k] = C::ldd(f2 + ((int64_t)k * g.S - dir_off(g, k)), bo);
      float e[Q7], phib[Q7];
#pragma unroll
      for (int k = 0; k < Q7; k++) {
        e[k] = C::ldd(h2 + ((int64_t)k * g.S - dir_off(g, k)), bo);
        phib[k] = v.src(k);
      }
      if (nonfluid) {
#pragma unroll
        for (int k = 1; k < Q27; k++)
          if ((nonfluid >> k) & 1u) raw[k] = C::ldd(f2 + (int64_t)opp27(k) * g.S, bo);
#pragma unroll
        for (int k = 1; k < Q7; k++)
          if ((nonfluid >> k) & 1u) {
            e[k] = C::ldd(h2 + (int64_t)opp7(k) * g.S, bo);
            phib[k] = v.c;
          }
      }
      float sumD, mx, my, mz;
      links27_dev<KIND>(
          g, i, [&](int k) { return raw[k]; }, [&](int k, float dk) { raw[k] = C::rtv(dk); },
          solid, sc.usx, sc.usy, sc.usz, rhoPhaseOld, exchange, x, y, z, cx, cy, cz, mom, sumD, mx,
          my, mz);
      finish_cell_dev<COLLIDE, KIND>(
          sc, i, near, phi_next, gxo, gyo, gzo, lapo, [&](int k) { return raw[k]; }, e, phib, sumD,
          mx, my, mz, use_bf, phi_next,
          [&](int k, float dv) { C::std_(fo + (int64_t)k * g.S, bo, dv); },
          [&](int k, float dv) { C::std_(ho + (int64_t)k * g.S, bo, dv); });
    } else {
      const Phi7 v = load_phi7(g, phi, bo4);
      const R phiOld = v.c;
      const R rhoPhaseOld = R(d.rhoG) + phiOld * R(d.dRho);
      float gxo, gyo, gzo, lapo;
      phi_grad7<R>(v, phiOld, gxo, gyo, gzo, lapo);
      uint32_t nonfluid, solid;
      if (near) tag_masks(g, i, tag, nonfluid, solid);
      else {
        nonfluid = wall_mask(g, x, y, z);
        solid = 0;
      }
      const RowAddr ra = row_addr(g, i, z * g.ny + y);
      float fp[Q27];
      R pSumF, mx, my, mz;
      pull27<KIND>(g, i, ra, f2, nonfluid, solid, sc.usx, sc.usy, sc.usz, rhoPhaseOld, exchange,
                   x, y, z, cx, cy, cz, mom, fp, pSumF, mx, my, mz);
      float hp[Q7];
      pull7<KIND>(g, ra, h2, nonfluid, hp);
      T *fob = fo + ra.base, *hob = ho + ra.base;
      finish_cell<COLLIDE, KIND>(
          sc, i, near, phi_next, gxo, gyo, gzo, lapo, fp, hp, pSumF, mx, my, mz, use_bf, phi_next,
          [&](int k, float val, float base) {
            C::sto(fob, ra.relb + (uint32_t)k * (uint32_t)g.pb * 4u, val, base);
          },
          [&](int k, float val, float base) {
            C::sto(hob, ra.relb + (uint32_t)k * (uint32_t)g.pb * 4u, val, base);
          });
    }
  }
  if (exchange) tile_momentum(mom, sc.partials, tile);
}
// Held to 128 VGPRs: four waves per SIMD; the band-row kernels and the
// one-scene kernel for every tile (the last step of a call) to 168 (three
// waves), where they fit without spilling.
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void km_stream_collide(
    const Scene *__restrict__ scenes, const int2 *__restrict__ tiles, int q, int first) {
  const int2 t = tiles[blockIdx.x];
  stream_collide_body<COLLIDE, KIND>(scenes[t.x], t.y, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k1_stream_collide(
    const Scene sc, int q, int first) {
  stream_collide_body<COLLIDE, KIND>(sc, blockIdx.x, q, first);
}
// The same over the bulk tiles and over the band-row tiles, one scene (k1) or
// a scene table (km: bulk over the tile table, band over each scene's band-tile
// launch, found by its prefix table off).
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void km_stream_collide_bulk(
    const Scene *__restrict__ scenes, const int2 *__restrict__ tiles, int q, int first) {
  const int2 t = tiles[blockIdx.x];
  stream_collide_body<COLLIDE, KIND, kTilesBulk>(scenes[t.x], t.y, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void km_stream_collide_band(
    const Scene *__restrict__ scenes, const int *__restrict__ off, int nscenes, int q, int first) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  stream_collide_body<COLLIDE, KIND, kTilesBand>(scenes[si], blockIdx.x - off[si], q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k1_stream_collide_bulk(
    const Scene sc, int q, int first) {
  stream_collide_body<COLLIDE, KIND, kTilesBulk>(sc, blockIdx.x, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k1_stream_collide_band(
    const Scene sc, int q, int first) {
  stream_collide_body<COLLIDE, KIND, kTilesBand>(sc, blockIdx.x, q, first);
}
// The band box's cells of the band-row tiles (beside the row march).
template <int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k1_stream_collide_box(
    const Scene sc, int q, int first) {
  stream_collide_body<true, KIND, kTilesBox>(sc, blockIdx.x, q, first);
}
// ---------------------------------------------------------------- stream + collide, two cells per lane

// The update of stream_collide_body for the 16-bit codecs with each lane owning
// two consecutive cells (i0, i0 + 1), i0 even: every population plane is read
// and written in naturally aligned 4-byte pairs instead of 2-byte elements,
// twice the bytes per memory instruction (the bare 34-plane pull pattern on
// MI355X: 5.2 against 3.8 TB/s).  A tile of 256 cells is one workgroup of two
// waves, so tiles, their order and the momentum partials are the one-cell
// kernel's.
//
// The source pair of a direction, (i0 - off, i0 + 1 - off), is an aligned pair
// when off is even.  When off is odd it is the upper element of the aligned pair
// at i0 - off - 1 and the lower element of the aligned pair after it; the lane
// loads both.  The parity of off follows from the parities of nx and
// nx ny (the parity class of the scene), so it is known per direction at
// compile time and all lanes take the same path.
//
// The two cells are updated one after the other.  Their pulled populations,
// and then their updated ones, stay packed in one register per direction, so
// the registers hold the working set of one cell at a time.  A pair holding a
// fluid cell is stored whole: the element of a non-fluid cell receives a zero
// deviation (populations of non-fluid cells are never read).

// 0: nx even; 1: nx odd, ny even; 2: nx and ny odd.
__host__ __device__ inline int parity_class(int nx, int ny) {
  return (nx & 1) == 0 ? 0 : ((ny & 1) == 0 ? 1 : 2);
}
// Parity of the pull offset ez nx ny + ey nx + ex of direction K in class PC
// (D3Q7 directions are D3Q27 directions 0..6).
template <int PC, int K>
struct PullOdd {
  static constexpr bool v =
      ((EX27[K] + (PC != 0 ? EY27[K] : 0) + (PC == 2 ? EZ27[K] : 0)) & 1) != 0;
};

// f(std::integral_constant<int, K>) for K = B .. E - 1.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Buffer resource over a population buffer of `bytes` bytes (raw addressing,
// no stride; word 3 as CDNA expects for untyped loads).  A load outside
// [0, bytes) returns 0 and a store there is dropped, which is how the pulls
// below stay in bounds without clamping: only lanes without a fluid cell
// address outside their buffer, and their values are never used.  Offsets are
// 32-bit byte offsets, so the buffers must be smaller than 4 GiB minus the
// largest pull offset (checked on the host).
constexpr int kRsrcWord3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kRsrcWord3);
}
// Cache policy of the pair loads and stores: stores are non-temporal (see stn);
// loads too in the one-scene form (NTL), where a grid's rows are long enough
// that the pulls' halo lines are not worth keeping (measured: 0.63 -> 0.60 ms at
// 256^3), but not for sets of small scenes, whose narrow rows share them.
constexpr int kCacheNT = 2;  // the nt bit of the cache-policy operand on gfx950
template <bool NTL = false>
__device__ __forceinline__ uint32_t ld_pair(__amdgpu_buffer_rsrc_t r, uint32_t e) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(2u * e), 0, NTL ? kCacheNT : 0);
}
__device__ __forceinline__ void st_pair(__amdgpu_buffer_rsrc_t r, uint32_t e, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)(2u * e), 0, kCacheNT);
}

// Elements p0 + i0 - off and p0 + i0 + 1 - off of a buffer of 16-bit elements
// (p0: the first element of the plane), packed, the first in the low half.
// Element indices wrap modulo 2^32 (a source before the buffer becomes an
// offset past its end).  For odd off the two elements straddle the aligned
// pairs (src - 1, src) and (src + 1, src + 2), and the lane loads both (lo,
// hi) and joins them with pull_join once every pull is in flight: no lane
// exchange, so the pulls of a cell are independent loads.  For a direction
// k >= 1 the lower pair's element index, k S + i0 - off - 1, is at least S - 2,
// so it does not wrap.
template <bool ODD, bool NTL>
__device__ __forceinline__ void pull_pair(__amdgpu_buffer_rsrc_t r, uint32_t p0, uint32_t i0,
                                          uint32_t off, uint32_t &lo, uint32_t &hi) {
  if constexpr (!ODD) {
    lo = ld_pair<NTL>(r, p0 + i0 - off);
  } else {
    const uint32_t e = p0 + i0 - off - 1u;
    lo = ld_pair<NTL>(r, e);
    hi = ld_pair<NTL>(r, e + 2u);
  }
}
template <bool ODD>
__device__ __forceinline__ uint32_t pull_join(uint32_t lo, uint32_t hi) {
  if constexpr (!ODD) return lo;
  else return (lo >> 16) | (hi << 16);
}

// A value stored into cell j (0: low half, 1: high half) of a packed pair; j is
// a constant once the loop over the pair's cells is unrolled.
template <typename C>
__device__ __forceinline__ void put_half(uint32_t &w, int j, float v) {
  if (j) C::template put<true>(w, v);
  else C::template put<false>(w, v);
}

// red: 6 x 256 doubles of workgroup memory for the tile's link momentum (not
// used by bulk tiles).  MODE as stream_collide_body: bulk tiles carry no tag
// reads and no momentum exchange.
template <bool COLLIDE, int KIND, int PC, bool NTL, int MODE>
__device__ __forceinline__ void stream_collide_pair_body(const Scene &sc, int tile, const int box[6],
                                                         int q, int first, double (*red)[256]) {
  using C = Codec<KIND>;
  using R = typename C::R;
  static_assert(KIND != ST_F32, "16-bit codecs only");
  const Geom &g = sc.g;
  const Derived &d = sc.d;
  const int64_t S = g.S;
  const uint32_t S32 = (uint32_t)S;
  const int lane = threadIdx.x;
  const int64_t wbase = ((int64_t)tile << kTileShift) + 128 * (int)threadIdx.y;
  const int64_t i0 = wbase + 2 * lane;
  const uint32_t i0u = (uint32_t)i0;
  q &= 1;
  const __amdgpu_buffer_rsrc_t f2 = pop_rsrc(par(sc.f, q), (uint32_t)(Q27 * 2 * S));
  const __amdgpu_buffer_rsrc_t h2 = pop_rsrc(par(sc.h, q), (uint32_t)(Q7 * 2 * S));
  const __amdgpu_buffer_rsrc_t fo = pop_rsrc(par(sc.f, q ^ 1), (uint32_t)(Q27 * 2 * S));
  const __amdgpu_buffer_rsrc_t ho = pop_rsrc(par(sc.h, q ^ 1), (uint32_t)(Q7 * 2 * S));
  const float *__restrict__ phi = par(sc.phi, q);
  float *__restrict__ phi_next = par(sc.phi, q ^ 1);
  const uint8_t *__restrict__ tag = sc.tag;
  const Ctrl *__restrict__ ctrl = sc.ctrl;
  const bool exchange = MODE != kTilesBulk && sc.has_solid != 0 &&
                        (MODE == kTilesBand || tile_in_box_rows(sc.tm, box, tile));
  const int use_bf = first && sc.use_bf;
  double cx = 0, cy = 0, cz = 0;
  if (exchange) {
    cx = ctrl->sphere.cx;
    cy = ctrl->sphere.cy;
    cz = ctrl->sphere.cz;
  }
  // Both cells: coordinates, fluid flags and source masks (neighbour tags
  // inside the band box, the walls elsewhere).
  int xs[2], ys[2], zs[2];
  bool fl[2], nr[2];
  uint32_t nf[2], so[2];
#pragma unroll
  for (int J = 0; J < 2; J++) {
    const int64_t i = i0 + J;
    cell_xyz(g.nx, g.nxny, i, xs[J], ys[J], zs[J]);
    nr[J] = MODE != kTilesBulk && in_box(box, xs[J], ys[J], zs[J]);
    fl[J] = i < g.n && interior(g, xs[J], ys[J], zs[J]) && (!nr[J] || tag[i] == TAG_FLUID);
    nf[J] = so[J] = 0;
    if (fl[J]) {
      if (nr[J]) tag_masks(g, i, tag, nf[J], so[J]);
      else nf[J] = wall_mask(g, xs[J], ys[J], zs[J]);
    }
  }
  // Pulled pairs: every load is issued before any is joined, so all of them
  // are in flight together (the loads take 54 registers at most).  Joined
  // inside each pull, the compiler waited for every odd pull's loads before
  // issuing the next direction's.
  uint32_t fr[Q27], hr[Q7], fx[Q27], hx[Q7];
  static_for<0, Q27>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    pull_pair<PullOdd<PC, k>::v, NTL>(f2, k * S32, i0u, (uint32_t)dir_off(g, k), fr[k], fx[k]);
  });
  static_for<0, Q7>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    pull_pair<PullOdd<PC, k>::v, NTL>(h2, k * S32, i0u, (uint32_t)dir_off(g, k), hr[k], hx[k]);
  });
  static_for<0, Q27>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    fr[k] = pull_join<PullOdd<PC, k>::v>(fr[k], fx[k]);
  });
  static_for<0, Q7>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    hr[k] = pull_join<PullOdd<PC, k>::v>(hr[k], hx[k]);
  });
  // Halfway bounce-back: the own opposite population where a source is not
  // fluid, requested for every such direction before any is merged.
  const uint32_t nfa = nf[0] | nf[1];
  if (nfa) {
    uint32_t ob[Q27], oh[Q7];
#pragma unroll
    for (int k = 1; k < Q27; k++)
      ob[k] = ((nfa >> k) & 1u) ? ld_pair<NTL>(f2, opp27(k) * S32 + i0u) : 0u;
#pragma unroll
    for (int k = 1; k < Q7; k++)
      oh[k] = ((nfa >> k) & 1u) ? ld_pair<NTL>(h2, opp7(k) * S32 + i0u) : 0u;
#pragma unroll
    for (int k = 1; k < Q27; k++) {
      const uint32_t bm = (((nf[0] >> k) & 1u) ? 0x0000ffffu : 0u) |
                          (((nf[1] >> k) & 1u) ? 0xffff0000u : 0u);
      fr[k] = (fr[k] & ~bm) | (ob[k] & bm);
    }
#pragma unroll
    for (int k = 1; k < Q7; k++) {
      const uint32_t bm = (((nf[0] >> k) & 1u) ? 0x0000ffffu : 0u) |
                          (((nf[1] >> k) & 1u) ? 0xffff0000u : 0u);
      hr[k] = (hr[k] & ~bm) | (oh[k] & bm);
    }
  }
  // Link momentum per cell, by cell of the tile (component-major), in red;
  // reduced below in the one-cell kernel's order.
#pragma unroll
  for (int J = 0; J < 2; J++) {
    const int64_t i = i0 + J;
      double mom[6] = {0, 0, 0, 0, 0, 0};
    if (fl[J]) {
      const int sh = 16 * J;
      const uint32_t nonfluid = nf[J], solid = so[J];
      const int x = xs[J], y = ys[J], z = zs[J];
      const Phi7 v = load_phi7(g, phi, (uint32_t)i * 4u);
      const float phiOld = v.c;
      const float rhoPhaseOld = sc.df.rhoG + phiOld * sc.df.dRho;
      float gxo, gyo, gzo, lapo;
      phi_grad7<float>(v, phiOld, gxo, gyo, gzo, lapo);
      // The deviations are decoded from the packed pairs where they are used:
      // for the sums here, and again for the collide; a corrected solid link
      // goes back into its pair as it would be stored.
      float sumD, mx, my, mz;
      links27_dev<KIND>(
          g, i, [&](int k) { return C::dv((uint16_t)(fr[k] >> sh)); },
          [&](int k, float dk) { put_half<C>(fr[k], J, dk); }, solid,
          sc.usx, sc.usy, sc.usz, rhoPhaseOld, exchange, x, y, z, cx, cy, cz, mom, sumD, mx, my,
          mz);
      // A pulled h has the phi of its source as its baseline, a bounced one the
      // phi of the cell.
      float e[Q7], phib[Q7];
#pragma unroll
      for (int k = 0; k < Q7; k++) {
        e[k] = C::dv((uint16_t)(hr[k] >> sh));
        phib[k] = ((nonfluid >> k) & 1u) ? v.c : v.src(k);
      }
      finish_cell_dev<COLLIDE, KIND>(
          sc, i, nr[J], phi_next, gxo, gyo, gzo, lapo,
          [&](int k) { return C::dv((uint16_t)(fr[k] >> sh)); }, e, phib, sumD, mx, my, mz, use_bf,
          phi_next,
          [&](int k, float dv) { put_half<C>(fr[k], J, dv); },
          [&](int k, float dv) { put_half<C>(hr[k], J, dv); });
    } else if (!COLLIDE && i < g.n) {
      phi_next[i] = phi[i];
    }
    if (exchange) {
      const int t = (int)(i - ((int64_t)tile << kTileShift));
#pragma unroll
      for (int c = 0; c < 6; c++) red[c][t] = mom[c];
    }
  }
  if (fl[0] || fl[1]) {
    if (!(fl[0] && fl[1])) {
      // The element of the non-fluid cell gets a zero deviation (its pulled
      // value may come from outside the grid).
      const uint32_t m = fl[0] ? 0x0000ffffu : 0xffff0000u;
#pragma unroll
      for (int k = 0; k < Q27; k++) fr[k] &= m;
#pragma unroll
      for (int k = 0; k < Q7; k++) hr[k] &= m;
    }
#pragma unroll
    for (int k = 0; k < Q27; k++) st_pair(fo, k * S32 + i0u, fr[k]);
#pragma unroll
    for (int k = 0; k < Q7; k++) st_pair(ho, k * S32 + i0u, hr[k]);
  }
  if (exchange) {
    // The one-cell kernel sums each wave of 64 cells by an xor butterfly with
    // offsets 32 .. 1 (every lane ends with the value lane 0 computes: at
    // offset o, cell j < o adds cell j + o to itself), then the four waves in
    // order.  The same tree, level by level, over the cells of each 64-cell
    // group gives the same sums bit for bit.
    const int tid = (int)threadIdx.y * 64 + lane;
    __syncthreads();
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      for (int w = tid; w < 24 * o; w += 128) {
        const int c = w / (4 * o), r = w - c * (4 * o), grp = r / o, j = r - grp * o;
        double *row = red[c] + grp * 64;
        row[j] = row[j] + row[j + o];
      }
      __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
      for (int c = 0; c < 6; c++)
        sc.partials[tile * 6 + c] = ((red[c][0] + red[c][64]) + red[c][128]) + red[c][192];
    }
  }
}

// Workgroups of 64 x 2 threads, one tile each.  The bulk kernels are held to at
// most 128 VGPRs so that four waves share a SIMD (measured on MI355X: 0.68
// against 0.72 ms at 3 waves for the 256^3 scene in FP16, one launch for all
// tiles then); the kernels that carry the band's paths (band-row tiles, and
// every tile of one launch) to 168, three waves, where they fit without
// spilling.  lb: the block's tile of a tile launch, or its index in a
// band-tile launch.
constexpr int kPairWaves = 2;
template <bool COLLIDE, int KIND, bool NTL, int MODE = kTilesAll>
__device__ __forceinline__ void stream_collide_pair(const Scene &sc, int lb, int q, int first) {
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = sc.ctrl->box_fused[j];
  int tile;
  if constexpr (MODE == kTilesBand || MODE == kTilesBox) {
    if (!band_tile(sc, box, lb, tile)) return;
  } else {
    if (!map_tile(sc.tm.T, sc.tm.T8, lb, tile)) return;
  }
  if constexpr (MODE == kTilesBulk) {
    if (tile_in_box_rows(sc.tm, box, tile)) return;
  }
  __shared__ double red[MODE == kTilesBulk ? 1 : 6][256];
  const int pc = parity_class(sc.g.nx, sc.g.ny);
  if (pc == 0) stream_collide_pair_body<COLLIDE, KIND, 0, NTL, MODE>(sc, tile, box, q, first, red);
  else if (pc == 1) stream_collide_pair_body<COLLIDE, KIND, 1, NTL, MODE>(sc, tile, box, q, first, red);
  else stream_collide_pair_body<COLLIDE, KIND, 2, NTL, MODE>(sc, tile, box, q, first, red);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void km_stream_collide2(
    const Scene *__restrict__ scenes, const int2 *__restrict__ tiles, int q, int first) {
  const int2 t = tiles[blockIdx.x];
  stream_collide_pair<COLLIDE, KIND, false>(scenes[t.x], t.y, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void k1_stream_collide2(const Scene sc, int q, int first) {
  stream_collide_pair<COLLIDE, KIND, true>(sc, blockIdx.x, q, first);
}
// Bulk tiles and band-row tiles in separate launches, one scene (k1) or a
// scene table (km: bulk over the tile table, band over each scene's band-tile
// launch, found by its prefix table off).
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) void k1_stream_collide2_bulk(
    const Scene sc, int q, int first) {
  stream_collide_pair<COLLIDE, KIND, true, kTilesBulk>(sc, blockIdx.x, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void k1_stream_collide2_band(
    const Scene sc, int q, int first) {
  stream_collide_pair<COLLIDE, KIND, true, kTilesBand>(sc, blockIdx.x, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) void km_stream_collide2_bulk(
    const Scene *__restrict__ scenes, const int2 *__restrict__ tiles, int q, int first) {
  const int2 t = tiles[blockIdx.x];
  stream_collide_pair<COLLIDE, KIND, false, kTilesBulk>(scenes[t.x], t.y, q, first);
}
template <bool COLLIDE, int KIND>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void km_stream_collide2_band(
    const Scene *__restrict__ scenes, const int *__restrict__ off, int nscenes, int q, int first) {
  const int si = find_scene(off, nscenes, blockIdx.x);
  stream_collide_pair<COLLIDE, KIND, false, kTilesBand>(scenes[si], blockIdx.x - off[si], q, first);
}

// phase_stream_body for two cells per lane (16-bit codecs), with the pair
// kernel's tiles, lanes and pulls: each direction's h pair is one aligned
// 4-byte load (two when the pull offset is odd) and each source phi pair two
// 4-byte loads, so a pair of cells takes about two thirds of the requests of
// two cells alone.  The sum is phase_sum_dev of the same deviations and
// baselines, so phi(t+1) is the one-cell kernel's bit for bit.  Loads go
// through buffer resources: an offset outside a buffer (a wall cell's source
// before the grid) reads 0 and is never used.
template <int KIND, int PC, bool NTL>
__device__ __forceinline__ void phase_stream_pair_body(const Scene &sc, int tile, int q) {
  using C = Codec<KIND>;
  static_assert(KIND != ST_F32, "16-bit codecs only");
  const Geom &g = sc.g;
  const int64_t S = g.S;
  const uint32_t S32 = (uint32_t)S;
  const int lane = threadIdx.x;
  const int64_t i0 = ((int64_t)tile << kTileShift) + 128 * (int)threadIdx.y + 2 * lane;
  const uint32_t i0u = (uint32_t)i0;
  q &= 1;
  const __amdgpu_buffer_rsrc_t h2 = pop_rsrc(par(sc.h, q), (uint32_t)(Q7 * 2 * S));
  const __amdgpu_buffer_rsrc_t pr = pop_rsrc(par(sc.phi, q), (uint32_t)(4 * S));
  float *__restrict__ phi_next = par(sc.phi, q ^ 1);
  const uint8_t *__restrict__ tag = sc.tag;
  int box[6];
#pragma unroll
  for (int j = 0; j < 6; j++) box[j] = sc.ctrl->box_fused[j];
  bool fl[2];
  uint32_t nf[2];
#pragma unroll
  for (int J = 0; J < 2; J++) {
    const int64_t i = i0 + J;
    int x, y, z;
    cell_xyz(g.nx, g.nxny, i, x, y, z);
    const bool near = in_box(box, x, y, z);
    fl[J] = i < g.n && interior(g, x, y, z) && (!near || ldb(tag, (uint32_t)i) == TAG_FLUID);
    nf[J] = 0;
    if (fl[J]) {
      if (near) {
#pragma unroll
        for (int k = 1; k < Q7; k++)
          nf[J] |= (uint32_t)(ldb(tag, (uint32_t)(i - dir_off(g, k))) != TAG_FLUID) << k;
      } else {
        nf[J] = wall_mask(g, x, y, z) & 0x7eu;
      }
    }
  }
  // Every load in flight before any is used.
  uint32_t hr[Q7], hx[Q7], ps[Q7][2];
  static_for<0, Q7>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const uint32_t off = (uint32_t)dir_off(g, k);
    pull_pair<PullOdd<PC, k>::v, NTL>(h2, k * S32, i0u, off, hr[k], hx[k]);
    ps[k][0] = __builtin_amdgcn_raw_buffer_load_b32(pr, (int)(4u * (i0u - off)), 0, 0);
    ps[k][1] = __builtin_amdgcn_raw_buffer_load_b32(pr, (int)(4u * (i0u + 1u - off)), 0, 0);
  });
  static_for<0, Q7>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    hr[k] = pull_join<PullOdd<PC, k>::v>(hr[k], hx[k]);
  });
  // Halfway bounce-back: the own opposite pair where a source is not fluid.
  const uint32_t nfa = nf[0] | nf[1];
  if (nfa) {
    uint32_t oh[Q7];
#pragma unroll
    for (int k = 1; k < Q7; k++)
      oh[k] = ((nfa >> k) & 1u) ? ld_pair<false>(h2, opp7(k) * S32 + i0u) : 0u;
#pragma unroll
    for (int k = 1; k < Q7; k++) {
      const uint32_t bm = (((nf[0] >> k) & 1u) ? 0x0000ffffu : 0u) |
                          (((nf[1] >> k) & 1u) ? 0xffff0000u : 0u);
      hr[k] = (hr[k] & ~bm) | (oh[k] & bm);
    }
  }
  float out[2];
#pragma unroll
  for (int J = 0; J < 2; J++) {
    // phi of the cell: the rest direction's source is the cell itself.
    const float pc = __builtin_bit_cast(float, ps[0][J]);
    if (fl[J]) {
      float e[Q7], phib[Q7], hfull[Q7];
#pragma unroll
      for (int k = 0; k < Q7; k++) {
        const bool bounce = (nf[J] >> k) & 1u;
        e[k] = C::dv((uint16_t)(hr[k] >> (16 * J)));
        phib[k] = bounce ? pc : __builtin_bit_cast(float, ps[k][J]);
      }
      out[J] = phase_sum_dev(e, phib, hfull);
    } else {
      out[J] = pc;
    }
  }
  if (i0 + 1 < g.n) {
    stb(phi_next, i0u * 4u, out[0]);
    stb(phi_next, i0u * 4u + 4u, out[1]);
  } else if (i0 < g.n) {
    stb(phi_next, i0u * 4u, out[0]);
  }
}
// NTL: non-temporal h loads, as for the pair stream kernel (one-scene launches).
template <int KIND, bool NTL>
__device__ __forceinline__ void phase_stream_pair(const Scene &sc, int lb, int q) {
  int tile;
  if (!map_tile(sc.tm.T, sc.tm.T8, lb, tile)) return;
  const int pc = parity_class(sc.g.nx, sc.g.ny);
  if (pc == 0) phase_stream_pair_body<KIND, 0, NTL>(sc, tile, q);
  else if (pc == 1) phase_stream_pair_body<KIND, 1, NTL>(sc, tile, q);
  else phase_stream_pair_body<KIND, 2, NTL>(sc, tile, q);
}
template <int KIND>
__global__ __launch_bounds__(128) void km_phase_stream2(const Scene *__restrict__ scenes,
                                                        const int2 *__restrict__ tiles, int q) {
  const int2 t = tiles[blockIdx.x];
  phase_stream_pair<KIND, false>(scenes[t.x], t.y, q);
}
template <int KIND>
__global__ __launch_bounds__(128) void k1_phase_stream2(const Scene sc, int q) {
  phase_stream_pair<KIND, true>(sc, blockIdx.x, q);
}

// ---------------------------------------------------------------- storage conversion

// Float32 populations (the reference's f and h) to a 16-bit codec, and back.
// The h baseline is phi of the cell (the phi the populations belong to).
// Float32 populations are row blocks, the 16-bit ones planes of stride S.
template <int KIND>
__global__ __launch_bounds__(256) void k_encode(Geom g, const float *__restrict__ f,
                                                const float *__restrict__ h,
                                                const float *__restrict__ phi,
                                                typename Codec<KIND>::T *__restrict__ cf,
                                                typename Codec<KIND>::T *__restrict__ ch) {
  using C = Codec<KIND>;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.n) return;
  const float pc = phi[i];
  const int64_t pe = pop_base(g, i), S = g.S;
#pragma unroll
  for (int k = 0; k < Q27; k++) C::st(cf, k * S + i, f[pe + k * g.pb], w27(k));
#pragma unroll
  for (int k = 0; k < Q7; k++) C::st(ch, k * S + i, h[pe + k * g.pb], w7(k) * pc);
}
template <int KIND>
__global__ __launch_bounds__(256) void k_decode(Geom g,
                                                const typename Codec<KIND>::T *__restrict__ cf,
                                                const typename Codec<KIND>::T *__restrict__ ch,
                                                const float *__restrict__ phi,
                                                float *__restrict__ f, float *__restrict__ h) {
  using C = Codec<KIND>;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.n) return;
  const float pc = phi[i];
  const int64_t pe = pop_base(g, i), S = g.S;
#pragma unroll
  for (int k = 0; k < Q27; k++) f[pe + k * g.pb] = C::ld(cf, k * S + i, w27(k));
#pragma unroll
  for (int k = 0; k < Q7; k++) h[pe + k * g.pb] = C::ld(ch, k * S + i, w7(k) * pc);
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

// ---------------------------------------------------------------- control blocks

// Copies the control block of every scene of a set (block j: scene j) into
// one array, so the host reads the set's blocks back with one transfer.
static_assert(sizeof(Ctrl) % 4 == 0, "control block copied in words");
__global__ __launch_bounds__(64) void km_gather_ctrl(const Scene *__restrict__ scenes,
                                                     Ctrl *__restrict__ out) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(scenes[blockIdx.x].ctrl);
  uint32_t *dst = reinterpret_cast<uint32_t *>(out + blockIdx.x);
  for (int w = threadIdx.x; w < (int)(sizeof(Ctrl) / 4); w += 64) dst[w] = src[w];
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
// Clears the fresh/dead lists of continued scenes (their last call kept the
// lists of its last retag readable) before their band retag.
__device__ __forceinline__ void reset_lists(const Scene &sc) {
  if (!sc.live) return;
  sc.ctrl->n_fresh = 0;
  sc.ctrl->n_dead = 0;
  sc.ctrl->n_solid = 0;
}
__global__ __launch_bounds__(64) void km_reset_lists(const Scene *__restrict__ scenes, int n) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j < n) reset_lists(scenes[j]);
}
__global__ __launch_bounds__(64) void k1_reset_lists(const Scene sc) {
  if (threadIdx.x == 0) reset_lists(sc);
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
