// Per-pass kernels: one kernel per pass of Sim.step with the reference's array
// semantics (f -> f2 -> f, h -> h2 -> h, stored gradients).  They back the
// per-pass C entry points (the drop-in for code that calls the passes one by
// one) and the generic-solid path.  Every cell's arithmetic is the reference's:
// IEEE double in the reference's operation order, float32 stores.
//
// Launch shape: blocks of 64x4 cells (one wave per x-row segment), grid
// (ceil(nx/64), ceil(ny/4), nz).
#pragma once

#include "codec.h"
#include "device_common.h"

namespace lbm {

// Element of population k of cell i: row blocks for float32 populations,
// planes of stride S for the 16-bit codecs.
template <int KIND>
__device__ __forceinline__ int64_t pop_elem(const Geom &g, int64_t i, int k) {
  if constexpr (KIND == ST_F32) return pop_base(g, i) + k * g.pb;
  else return (int64_t)k * g.S + i;
}

#define LBM_CELL_XYZ                                 \
  const int x = blockIdx.x * 64 + threadIdx.x;       \
  const int y = blockIdx.y * 4 + threadIdx.y;        \
  const int z = blockIdx.z;                          \
  if (x >= g.nx || y >= g.ny) return;

// initFlat body (src/sim.js:112-132): phi from the per-column tanh profile
// (phicol holds the double phiCol, per y when per_cell == 0, per cell otherwise),
// rho = 1, u = 0, walls on box faces, f = w, h = w7 * phiCol.
__global__ __launch_bounds__(256) void k_init_cells(Geom g, const double *__restrict__ phicol,
                                                    int per_cell, float *__restrict__ phi,
                                                    float *__restrict__ phiPrev, float *rho,
                                                    float *ux, float *uy, float *uz, uint8_t *tag,
                                                    uint8_t *tagPrev, float *f, float *h) {
  LBM_CELL_XYZ
  const int64_t i = cell_index(g, x, y, z);
  const double pc = per_cell ? phicol[i] : phicol[y];
  phi[i] = (float)pc;
  phiPrev[i] = (float)pc;
  rho[i] = 1.0f;
  ux[i] = 0.0f;
  uy[i] = 0.0f;
  uz[i] = 0.0f;
  const bool onWall = x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1;
  tag[i] = onWall ? TAG_WALL : TAG_FLUID;
  tagPrev[i] = tag[i];
  const int64_t pe = pop_base(g, i);
#pragma unroll
  for (int k = 0; k < Q27; k++) f[pe + k * g.pb] = w27(k);
#pragma unroll
  for (int k = 0; k < Q7; k++) h[pe + k * g.pb] = (float)((double)w7(k) * pc);
}

// seedHydrostatic (src/sim.js:144-173): one thread per (x,z) column, sequential
// prefix along +y.
__global__ __launch_bounds__(256) void k_seed_hydrostatic(Geom g, double rhoG, double dRho,
                                                          double gravity,
                                                          const float *__restrict__ phi,
                                                          float *__restrict__ rho,
                                                          float *__restrict__ f) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int z = blockIdx.y;
  if (x >= g.nx) return;
  const double gOverCs2 = gravity * INV_CS2_27;
  const double uLatY = -0.5 * gravity;
  const double uLat2 = uLatY * uLatY;
  double pCs2 = 1.0;
  for (int y = 0; y < g.ny; y++) {
    const int64_t i = cell_index(g, x, y, z);
    const int64_t pe = pop_base(g, i);
    rho[i] = (float)pCs2;
    const double rhoPhase = rhoG + (double)phi[i] * dRho;
#pragma unroll
    for (int k = 0; k < Q27; k++) {
      const double eu = (double)EY27[k] * uLatY;
      const double Qi = eu * (INV_CS2_27 + INV_2CS4_27 * eu) - 1.5 * uLat2;
      f[pe + k * g.pb] = (float)((double)w27(k) * (pCs2 + rhoPhase * Qi));
    }
    pCs2 += rhoPhase * gOverCs2;
  }
}

// Appends index i to a device list; overflow sets kErrListOverflow.
__device__ __forceinline__ void list_push(int *counter, int32_t *list, int cap, int32_t i,
                                          int *err) {
  const int slot = atomicAdd(counter, 1);
  if (slot < cap) list[slot] = i;
  else atomicOr(err, kErrListOverflow);
}

// retagCells (src/sim.js:214-244) for one cell.
// mode 0: no solid; 1: the sphere b; 2: host mask + surface velocities.
// copy_prev: also snapshot phi into phiPrev (the per-pass retag does; the fused
// step does not, because there phiPrev is the other phi buffer).
// Cells that turn fluid (fresh) or solid (dead) are appended to the lists, and
// solid cells are counted.
__device__ __forceinline__ void retag_cell(const Geom &g, int mode, const lbm_sphere &b,
                                           const uint8_t *__restrict__ mask,
                                           const float *__restrict__ mus, uint8_t *tag,
                                           uint8_t *tagPrev, const float *__restrict__ phi,
                                           float *__restrict__ phiPrev, int copy_prev, float *usx,
                                           float *usy, float *usz, Ctrl *ctrl, int32_t *fresh,
                                           int32_t *dead, int cap, int x, int y, int z) {
  const int64_t i = cell_index(g, x, y, z);
  const uint8_t tp = tag[i];
  tagPrev[i] = tp;
  if (copy_prev) phiPrev[i] = phi[i];
  uint8_t t;
  float vx = 0.0f, vy = 0.0f, vz = 0.0f;
  if (!interior(g, x, y, z)) {
    t = TAG_WALL;
  } else {
    bool solid = false;
    if (mode == 1) {
      solid = sphere_contains(b, x, y, z);
      if (solid) {
        double tmp[3];
        sphere_surface_velocity(b, x, y, z, tmp);
        vx = (float)tmp[0];
        vy = (float)tmp[1];
        vz = (float)tmp[2];
      }
    } else if (mode == 2) {
      solid = mask[i] != 0;
      if (solid) {
        vx = mus[3 * i + 0];
        vy = mus[3 * i + 1];
        vz = mus[3 * i + 2];
      }
    }
    t = solid ? TAG_SOLID : TAG_FLUID;
    if (t == TAG_FLUID && tp == TAG_SOLID) list_push(&ctrl->n_fresh, fresh, cap, (int32_t)i, &ctrl->err);
    if (t == TAG_SOLID && tp == TAG_FLUID) list_push(&ctrl->n_dead, dead, cap, (int32_t)i, &ctrl->err);
  }
  tag[i] = t;
  usx[i] = vx;
  usy[i] = vy;
  usz[i] = vz;
}

// retagCells over the whole grid (per-pass API).
__global__ __launch_bounds__(256) void k_retag(Geom g, int mode, const lbm_sphere *__restrict__ sph,
                                               const uint8_t *__restrict__ mask,
                                               const float *__restrict__ mus, uint8_t *tag,
                                               uint8_t *tagPrev, const float *__restrict__ phi,
                                               float *__restrict__ phiPrev, int copy_prev,
                                               float *usx, float *usy, float *usz, Ctrl *ctrl,
                                               int32_t *fresh, int32_t *dead, int cap) {
  LBM_CELL_XYZ
  lbm_sphere b{};
  if (mode == 1) b = *sph;
  retag_cell(g, mode, b, mask, mus, tag, tagPrev, phi, phiPrev, copy_prev, usx, usy, usz, ctrl,
             fresh, dead, cap, x, y, z);
}

// Sorts a[0..count) ascending within one workgroup (bitonic).  Lists up to
// kSortLds entries sort in the LDS buffer sh; longer ones sort in place in
// global memory, padded with INT_MAX up to the next power of two (list
// capacities are powers of two).
constexpr int kSortLds = 8192;
__device__ void sort_list_block(int32_t *a, int count, int32_t *sh) {
  if (count <= 1) return;
  int P = 1;
  while (P < count) P <<= 1;
  const int tid = threadIdx.x;
  int32_t *w;
  if (P <= kSortLds) {
    for (int t = tid; t < P; t += blockDim.x) sh[t] = t < count ? a[t] : INT32_MAX;
    w = sh;
  } else {
    for (int t = count + tid; t < P; t += blockDim.x) a[t] = INT32_MAX;
    w = a;
  }
  __threadfence_block();
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < P; t += blockDim.x) {
        const int u = t ^ j;
        if (u > t) {
          const int32_t at = w[t], au = w[u];
          const bool up = (t & k) == 0;
          if ((at > au) == up) {
            w[t] = au;
            w[u] = at;
          }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  }
  if (P <= kSortLds)
    for (int t = tid; t < count; t += blockDim.x) a[t] = sh[t];
  __threadfence_block();
  __syncthreads();
}

// Sorts the fresh (block 0) or dead (block 1) list of the last retag.
__global__ __launch_bounds__(1024) void k_sort_lists(int32_t *fresh, int32_t *dead, Ctrl *ctrl,
                                                     int cap) {
  __shared__ int32_t sh[kSortLds];
  int count = blockIdx.x == 0 ? ctrl->n_fresh : ctrl->n_dead;
  if (count > cap) count = cap;
  sort_list_block(blockIdx.x == 0 ? fresh : dead, count, sh);
}

__device__ __forceinline__ int lower_bound_i32(const int32_t *a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Neighbour n = 0..25 of a cell in the reference's visiting order (dz, dy, dx
// each from -1 to 1, the cell itself skipped): its cell index offset.  The
// first 13 are the neighbours with a smaller index.
__device__ __forceinline__ int64_t nb_offset(const Geom &g, int n) {
  const int m = n < 13 ? n : n + 1;
  return (int64_t)(m / 9 - 1) * g.nxny + (int64_t)((m / 3) % 3 - 1) * g.nx + (m % 3 - 1);
}

// handleFreshCells (src/sim.js:257-310) over the sorted fresh list, in one
// workgroup.  The reference updates cells in place in index order, so a fresh
// cell averages the already-reseeded values of earlier fresh neighbours and the
// stale values of later ones.  Cells are scheduled by dependency level: a cell's
// level is one more than the highest level of its earlier fresh neighbours, and
// all cells of a level are processed together, which reproduces the sequential
// order exactly (a later fresh neighbour always has a higher level).
//
// phi_fb: phiPrev read for the no-neighbour fallback; phi_w1/phi_w2: the two
// writes of phi[i] = phiPrev[i] = p (they may alias); neighbours read phi_w1.
// f and h receive the reseeded populations.  Inlined into its kernels: as a
// called function its arguments and the caller's live registers went through
// scratch memory.
template <int KIND>
__device__ __forceinline__ void fresh_block(Geom g, double rhoG, double rhoL,
                                            const int32_t *__restrict__ list,
                            int count, int *level_scratch, const uint8_t *__restrict__ tag,
                            const float *phi_fb, float *phi_w1, float *phi_w2, float *rho,
                            float *ux, float *uy, float *uz, typename Codec<KIND>::T *f,
                            typename Codec<KIND>::T *h) {
  using C = Codec<KIND>;
  constexpr int kLds = 4096;
  __shared__ int sh_level[kLds];
  __shared__ int32_t sh_list[kLds];
  __shared__ int sh_flag;
  if (count == 0) return;
  const int tid = threadIdx.x;
  int *lv = count <= kLds ? sh_level : level_scratch;
  const int32_t *sl = list;
  if (count <= kLds) {
    for (int c = tid; c < count; c += blockDim.x) sh_list[c] = list[c];
    sl = sh_list;
  }
  for (int c = tid; c < count; c += blockDim.x) lv[c] = 0;
  __threadfence_block();
  __syncthreads();
  // Longest-path levels by repeated relaxation (terminates after depth+1
  // rounds).  The 13 earlier neighbours' tags are loaded together, so a round
  // waits for memory once per cell.
  for (;;) {
    if (tid == 0) sh_flag = 0;
    __syncthreads();
    for (int c = tid; c < count; c += blockDim.x) {
      const int32_t i = sl[c];
      uint8_t tg[13];
#pragma unroll
      for (int n = 0; n < 13; n++) tg[n] = tag[i + nb_offset(g, n)];
      int L = 0;
#pragma unroll
      for (int n = 0; n < 13; n++) {
        if (tg[n] != TAG_FLUID) continue;
        const int32_t j = (int32_t)(i + nb_offset(g, n));
        const int q = lower_bound_i32(sl, c, j);
        if (q < c && sl[q] == j) {
          const int lq = lv[q] + 1;
          if (lq > L) L = lq;
        }
      }
      if (L != lv[c]) {
        lv[c] = L;
        sh_flag = 1;
      }
    }
    __threadfence_block();
    __syncthreads();
    if (!sh_flag) break;
    __syncthreads();
  }
  if (tid == 0) sh_flag = 0;
  __syncthreads();
  for (int c = tid; c < count; c += blockDim.x)
    if (lv[c]) atomicMax(&sh_flag, lv[c]);
  __syncthreads();
  const int maxL = sh_flag;
  const double dRho = rhoL - rhoG;
  for (int L = 0; L <= maxL; L++) {
    for (int c = tid; c < count; c += blockDim.x) {
      if (lv[c] != L) continue;
      const int32_t i = sl[c];
      double psum = 0, rsum = 0, vxs = 0, vys = 0, vzs = 0;
      int pcnt = 0;
      // The neighbours one z plane at a time (9, 8 and 9 cells): their tags
      // and fields are loaded together, then summed in the visiting order.
#pragma unroll
      for (int grp = 0; grp < 3; grp++) {
        const int n0 = grp == 0 ? 0 : grp == 1 ? 9 : 17;
        const int n1 = grp == 0 ? 9 : grp == 1 ? 17 : 26;
        uint8_t tg[9];
        float pv[9], rv[9], xv[9], yv[9], zv[9];
#pragma unroll
        for (int u = 0; u < 9; u++) {
          if (n0 + u >= n1) continue;
          const int64_t j = i + nb_offset(g, n0 + u);
          tg[u] = tag[j];
          pv[u] = phi_w1[j];
          rv[u] = rho[j];
          xv[u] = ux[j];
          yv[u] = uy[j];
          zv[u] = uz[j];
        }
#pragma unroll
        for (int u = 0; u < 9; u++) {
          if (n0 + u >= n1 || tg[u] != TAG_FLUID) continue;
          psum += (double)pv[u];
          rsum += (double)rv[u];
          vxs += (double)xv[u];
          vys += (double)yv[u];
          vzs += (double)zv[u];
          pcnt++;
        }
      }
      const double p = pcnt > 0 ? psum / pcnt : (double)phi_fb[i];
      const double r = pcnt > 0 ? rsum / pcnt : 1.0;
      const double vx = pcnt > 0 ? vxs / pcnt : 0;
      const double vy = pcnt > 0 ? vys / pcnt : 0;
      const double vz = pcnt > 0 ? vzs / pcnt : 0;
      const float pf = (float)p;
      phi_w1[i] = pf;
      phi_w2[i] = pf;
      rho[i] = (float)r;
      ux[i] = (float)vx;
      uy[i] = (float)vy;
      uz[i] = (float)vz;
      const double rhoPhase = rhoG + p * dRho;
      const double u2 = vx * vx + vy * vy + vz * vz;
#pragma unroll
      for (int k = 0; k < Q27; k++) {
        const double eu = EX27[k] * vx + EY27[k] * vy + EZ27[k] * vz;
        const double Qi = eu * (INV_CS2_27 + INV_2CS4_27 * eu) - 1.5 * u2;
        C::st(f, pop_elem<KIND>(g, i, k), (float)((double)w27(k) * (r + rhoPhase * Qi)), w27(k));
      }
#pragma unroll
      for (int k = 0; k < Q7; k++) {
        const double eu = EX7[k] * vx + EY7[k] * vy + EZ7[k] * vz;
        C::st(h, pop_elem<KIND>(g, i, k), (float)((double)w7(k) * p * (1 + INV_CS2_7 * eu)),
              w7(k) * pf);
      }
    }
    // The next level reads these cells from other threads of this workgroup
    // only: the barrier's workgroup-scope fence makes the writes visible (an
    // agent-scope fence would write back L2 at every level).
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void k_fresh(Geom g, double rhoG, double rhoL,
                                                const int32_t *__restrict__ list, Ctrl *ctrl,
                                                int cap, int *level_scratch,
                                                const uint8_t *__restrict__ tag,
                                                const float *phi_fb, float *phi_w1, float *phi_w2,
                                                float *rho, float *ux, float *uy, float *uz,
                                                float *f, float *h) {
  int count = ctrl->n_fresh;
  if (count > cap) count = cap;
  fresh_block<ST_F32>(g, rhoG, rhoL, list, count, level_scratch, tag, phi_fb, phi_w1, phi_w2, rho,
                      ux, uy, uz, f, h);
}

// computePhiGradients (src/sim.js:176-199): all cells, clamped indices.
__global__ __launch_bounds__(256) void k_gradients(Geom g, const float *__restrict__ phi,
                                                   float *__restrict__ gx, float *__restrict__ gy,
                                                   float *__restrict__ gz,
                                                   float *__restrict__ lap) {
  LBM_CELL_XYZ
  const int xm = x > 0 ? x - 1 : x, xp = x < g.nx - 1 ? x + 1 : x;
  const int ym = y > 0 ? y - 1 : y, yp = y < g.ny - 1 ? y + 1 : y;
  const int zm = z > 0 ? z - 1 : z, zp = z < g.nz - 1 ? z + 1 : z;
  const int64_t i = cell_index(g, x, y, z);
  const double c = phi[i];
  const double px = phi[cell_index(g, xp, y, z)];
  const double mx = phi[cell_index(g, xm, y, z)];
  const double py = phi[cell_index(g, x, yp, z)];
  const double my = phi[cell_index(g, x, ym, z)];
  const double pz = phi[cell_index(g, x, y, zp)];
  const double mz = phi[cell_index(g, x, y, zm)];
  gx[i] = (float)(0.5 * (px - mx));
  gy[i] = (float)(0.5 * (py - my));
  gz[i] = (float)(0.5 * (pz - mz));
  lap[i] = (float)(px + mx + py + my + pz + mz - 6 * c);
}

// collidePhase (src/sim.js:324-364).
__global__ __launch_bounds__(256) void k_collide_phase(Geom g, Derived d,
                                                       const float *__restrict__ h,
                                                       float *__restrict__ h2,
                                                       const float *__restrict__ phi,
                                                       const float *__restrict__ ux,
                                                       const float *__restrict__ uy,
                                                       const float *__restrict__ uz,
                                                       const uint8_t *__restrict__ tag,
                                                       const float *__restrict__ gx,
                                                       const float *__restrict__ gy,
                                                       const float *__restrict__ gz) {
  LBM_CELL_XYZ
  if (!interior(g, x, y, z)) return;
  const int64_t i = cell_index(g, x, y, z);
  if (tag[i] != TAG_FLUID) return;
  const double p = phi[i];
  const double uxi = ux[i], uyi = uy[i], uzi = uz[i];
  const double ggx = gx[i], ggy = gy[i], ggz = gz[i];
  const double gmag = sqrt(ggx * ggx + ggy * ggy + ggz * ggz) + 1e-12;
  const double sharpen = d.sharpC * p * (1 - p);
  const double Sx = sharpen * d.sharpFac * ggx / gmag;
  const double Sy = sharpen * d.sharpFac * ggy / gmag;
  const double Sz = sharpen * d.sharpFac * ggz / gmag;
  const int64_t pe = pop_base(g, i);
#pragma unroll
  for (int k = 0; k < Q7; k++) {
    const double ekx = EX7[k], eky = EY7[k], ekz = EZ7[k];
    const double wk = w7(k);
    const double eu = ekx * uxi + eky * uyi + ekz * uzi;
    const double heq = wk * p * (1 + INV_CS2_7 * eu);
    const double Fk = wk * (ekx * Sx + eky * Sy + ekz * Sz);
    const double hk = h[pe + k * g.pb];
    h2[pe + k * g.pb] = (float)(hk - d.invTauPhi * (hk - heq) + Fk);
  }
}

// streamPhase (src/sim.js:376-397): pull; a non-fluid source bounces back.
__global__ __launch_bounds__(256) void k_stream_phase(Geom g, const float *__restrict__ h2,
                                                      float *__restrict__ h,
                                                      const uint8_t *__restrict__ tag) {
  LBM_CELL_XYZ
  if (!interior(g, x, y, z)) return;
  const int64_t i = cell_index(g, x, y, z);
  if (tag[i] != TAG_FLUID) return;
  const int64_t pe = pop_base(g, i);
#pragma unroll
  for (int k = 0; k < Q7; k++) {
    const int64_t s = i - ((int64_t)EZ7[k] * g.nxny + (int64_t)EY7[k] * g.nx + EX7[k]);
    h[pe + k * g.pb] = tag[s] == TAG_FLUID ? h2[pop_base(g, s) + k * g.pb]
                                                    : h2[pe + opp7(k) * g.pb];
  }
}

// collideHydro (src/sim.js:406-464).
__global__ __launch_bounds__(256) void k_collide_hydro(
    Geom g, Derived d, const float *__restrict__ f, float *__restrict__ f2,
    const float *__restrict__ rho, const float *__restrict__ ux, const float *__restrict__ uy,
    const float *__restrict__ uz, const float *__restrict__ phi, const uint8_t *__restrict__ tag,
    const float *__restrict__ gx, const float *__restrict__ gy, const float *__restrict__ gz,
    const float *__restrict__ lap, const float *__restrict__ bfx, const float *__restrict__ bfy,
    const float *__restrict__ bfz, int use_bf) {
  LBM_CELL_XYZ
  if (!interior(g, x, y, z)) return;
  const int64_t i = cell_index(g, x, y, z);
  if (tag[i] != TAG_FLUID) return;
  const double sumF = rho[i];
  const double p = phi[i];
  const double rhoPhase = d.rhoG + p * d.dRho;
  const double mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * (double)lap[i];
  const double bx = use_bf ? (double)bfx[i] : 0.0;
  const double by = use_bf ? (double)bfy[i] : 0.0;
  const double bz = use_bf ? (double)bfz[i] : 0.0;
  const double fx = bx + mu * (double)gx[i];
  const double fy = by + mu * (double)gy[i] + rhoPhase * d.gravity;
  const double fz = bz + mu * (double)gz[i];
  const double uxe = ux[i], uye = uy[i], uze = uz[i];
  const double u2 = uxe * uxe + uye * uye + uze * uze;
  const double negU2half = -1.5 * u2;
  const double uF = uxe * fx + uye * fy + uze * fz;
  const int64_t pe = pop_base(g, i);
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    const double ekx = EX27[k], eky = EY27[k], ekz = EZ27[k];
    const double wk = w27(k);
    const double eu = ekx * uxe + eky * uye + ekz * uze;
    const double Qi = negU2half + eu * (INV_CS2_27 + INV_2CS4_27 * eu);
    const double feq = wk * (sumF + rhoPhase * Qi);
    const double ekF = ekx * fx + eky * fy + ekz * fz;
    const double Sk = d.A * wk * (ekF * (1 + INV_CS2_27 * eu) - uF);
    const double fk = f[pe + k * g.pb];
    f2[pe + k * g.pb] = (float)(fk - d.invTau * (fk - feq) + Sk);
  }
}

// streamHydroAndMacro (src/sim.js:484-582).  With emit_links, each fluid->solid
// link appends (key = i*27 + k, m = f_opp + f_k) in no particular order; sorted
// by key, the list gives the applyImpulseAtCell calls in the reference's order.
__global__ __launch_bounds__(256) void k_stream_hydro_macro(
    Geom g, Derived d, const float *__restrict__ f2, float *__restrict__ f,
    const float *__restrict__ h, float *__restrict__ rho, float *__restrict__ ux,
    float *__restrict__ uy, float *__restrict__ uz, float *__restrict__ phi,
    const uint8_t *__restrict__ tag, const float *__restrict__ usx, const float *__restrict__ usy,
    const float *__restrict__ usz, const float *__restrict__ gx, const float *__restrict__ gy,
    const float *__restrict__ gz, const float *__restrict__ lap, const float *__restrict__ bfx,
    const float *__restrict__ bfy, const float *__restrict__ bfz, int use_bf, int emit_links,
    int64_t *link_key, double *link_m, int link_cap, Ctrl *ctrl) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  bool active = x < g.nx && y < g.ny && interior(g, x, y, z);
  const int64_t i = active ? cell_index(g, x, y, z) : 0;
  active = active && tag[i] == TAG_FLUID;
  int slot = 0;
  if (emit_links) {
    // One list reservation per workgroup: each lane counts its links (solid
    // neighbours), takes its offset in LDS, and one thread reserves the
    // total (one global atomic per workgroup instead of one per link).
    __shared__ int blk_n, blk_base;
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    if (lead) blk_n = 0;
    __syncthreads();
    int nl = 0;
    if (active) {
#pragma unroll
      for (int k = 1; k < Q27; k++)
        nl += tag[i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k])] == TAG_SOLID;
    }
    if (nl) slot = atomicAdd(&blk_n, nl);
    __syncthreads();
    if (lead && blk_n) blk_base = atomicAdd(&ctrl->n_links, blk_n);
    __syncthreads();
    if (nl) slot += blk_base;
  }
  if (!active) return;
  const double rhoPhaseOld = d.rhoG + (double)phi[i] * d.dRho;
  const int64_t pe = pop_base(g, i);
  double pSumF = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int k = 0; k < Q27; k++) {
    const int64_t s = i - ((int64_t)EZ27[k] * g.nxny + (int64_t)EY27[k] * g.nx + EX27[k]);
    const uint8_t st = tag[s];
    const double ekx = EX27[k], eky = EY27[k], ekz = EZ27[k];
    double fk;
    if (st == TAG_FLUID) {
      fk = f2[pop_base(g, s) + k * g.pb];
    } else {
      const double fopp = f2[pe + opp27(k) * g.pb];
      double corr = 0;
      if (st == TAG_SOLID) {
        const double eu = ekx * (double)usx[s] + eky * (double)usy[s] + ekz * (double)usz[s];
        corr = 2 * (double)w27(k) * rhoPhaseOld * eu * INV_CS2_27;
      }
      fk = fopp + corr;
      if (st == TAG_SOLID && emit_links) {
        const double m = fopp + fk;
        if (slot < link_cap) {
          link_key[slot] = i * Q27 + k;
          link_m[slot] = m;
        } else {
          atomicOr(&ctrl->err, kErrListOverflow);
        }
        slot++;
      }
    }
    f[pe + k * g.pb] = (float)fk;
    pSumF += fk;
    mx += ekx * fk;
    my += eky * fk;
    mz += ekz * fk;
  }
  rho[i] = (float)(pSumF > 0.01 ? pSumF : 0.01);
  double phiSum = 0;
#pragma unroll
  for (int k = 0; k < Q7; k++) phiSum += (double)h[pe + k * g.pb];
  const double p = phiSum < 0 ? 0 : phiSum > 1 ? 1 : phiSum;
  phi[i] = (float)p;
  const double rhoPhaseNew = d.rhoG + p * d.dRho;
  const double mu = 4 * d.beta * p * (p - 1) * (2 * p - 1) - d.kappa * (double)lap[i];
  const double bx = use_bf ? (double)bfx[i] : 0.0;
  const double by = use_bf ? (double)bfy[i] : 0.0;
  const double bz = use_bf ? (double)bfz[i] : 0.0;
  const double fx = bx + mu * (double)gx[i];
  const double fy = by + mu * (double)gy[i] + rhoPhaseNew * d.gravity;
  const double fz = bz + mu * (double)gz[i];
  const double invRho = 1 / rhoPhaseNew;
  double uxn = (mx + 0.5 * fx) * invRho;
  double uyn = (my + 0.5 * fy) * invRho;
  double uzn = (mz + 0.5 * fz) * invRho;
  const double u2 = uxn * uxn + uyn * uyn + uzn * uzn;
  const double UMAX = 0.2, UMAX2 = UMAX * UMAX;
  if (u2 > UMAX2) {
    const double sc = UMAX / sqrt(u2);
    uxn *= sc;
    uyn *= sc;
    uzn *= sc;
  }
  ux[i] = (float)uxn;
  uy[i] = (float)uyn;
  uz[i] = (float)uzn;
}

// Counts solid cells into ctrl->n_solid (bounds the momentum-exchange link list:
// every link has a solid source cell, and each cell sources at most 26 links).
__global__ __launch_bounds__(256) void k_count_solid(int64_t n, const uint8_t *__restrict__ tag,
                                                     Ctrl *ctrl) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int local = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    local += tag[i] == TAG_SOLID;
  if (local) atomicAdd(&cnt, local);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&ctrl->n_solid, cnt);
}

// ---------------------------------------------------------------- sphere impulses on the device

// The per-pass step's momentum exchange and dead-cell impulses applied to the
// device sphere (ctrl->sphere) in the reference's order, so the host gets the
// sphere back in one transfer instead of the link list.  Per chunk of 1024
// records, every thread computes one record's six increments (impulse_terms,
// the host's function, FP64) into LDS; then lanes 0..5 of the first wave add
// them to Fx, Fy, Fz, Tx, Ty, Tz record after record.  Each sum sees the same
// additions in the same order as applyImpulseAtCell called once per record
// (src/solid.js:70-80), so the sphere ends bit for bit as the host's replay.

constexpr int kApplyChunk = 1024;

// rec(j, x, y, z, px, py, pz) yields record j; cnt records in order.
template <class Rec>
__device__ __forceinline__ void apply_impulses(Ctrl *ctrl, int cnt, const Rec &rec) {
  __shared__ double terms[kApplyChunk][6];
  if (cnt <= 0) return;
  const int tid = threadIdx.x;
  const lbm_sphere b = ctrl->sphere;
  double *sums = &ctrl->sphere.Fx;  // Fx, Fy, Fz, Tx, Ty, Tz are consecutive
  const int q = tid < 6 ? tid : 0;
  double acc = sums[q];
  for (int base = 0; base < cnt; base += kApplyChunk) {
    const int j = base + tid;
    if (j < cnt) {
      int x, y, z;
      double px, py, pz;
      rec(j, x, y, z, px, py, pz);
      impulse_terms(b, x, y, z, px, py, pz, terms[tid]);
    }
    __syncthreads();
    if (tid < 64) {
      // One dependent add per record; the loads of the next 16 records are
      // issued before the adds of the current 16 so LDS latency stays hidden.
      const int e = cnt - base < kApplyChunk ? cnt - base : kApplyChunk;
      constexpr int U = 16;
      int t = 0;
      if (e >= U) {
        double cur[U];
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = terms[u][q];
        for (; t + 2 * U <= e; t += U) {
          double nxt[U];
#pragma unroll
          for (int u = 0; u < U; u++) nxt[u] = terms[t + U + u][q];
#pragma unroll
          for (int u = 0; u < U; u++) acc += cur[u];
#pragma unroll
          for (int u = 0; u < U; u++) cur[u] = nxt[u];
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc += cur[u];
        t += U;
      }
      for (; t < e; t++) acc += terms[t][q];
    }
    __syncthreads();
  }
  if (tid < 6) sums[tid] = acc;
}

// Link (i, k) with momentum m (src/sim.js:530-535) is applyImpulseAtCell(
// x - ex, y - ey, z - ez, -m ex, -m ey, -m ez) for cell i = (x, y, z).
__device__ __forceinline__ void link_impulse(const Geom &g, int64_t key, double m, int &x, int &y,
                                             int &z, double &px, double &py, double &pz) {
  const int64_t i = key / Q27;
  const int k = (int)(key - i * Q27);
  const int ex = dir_comp(kEX27Bits, k), ey = dir_comp(kEY27Bits, k), ez = dir_comp(kEZ27Bits, k);
  x = (int)(i % g.nx) - ex;
  y = (int)((i / g.nx) % g.ny) - ey;
  z = (int)(i / g.nxny) - ez;
  px = -m * (double)ex;
  py = -m * (double)ey;
  pz = -m * (double)ez;
}

// Links sorted by key = cell * 27 + k: cells in index order, directions
// k = 0..26 within a cell, the reference's loop order (src/sim.js:509-543).
struct SortedLinkRecord {
  Geom g;
  const int64_t *keys;
  const double *ms;
  __device__ void operator()(int j, int &x, int &y, int &z, double &px, double &py,
                             double &pz) const {
    link_impulse(g, keys[j], ms[j], x, y, z, px, py, pz);
  }
};

// Lists of sorted links (rocprim's radix sort ran first).
__global__ __launch_bounds__(kApplyChunk) void k_apply_links(Geom g, const int64_t *__restrict__ keys,
                                                             const double *__restrict__ ms,
                                                             Ctrl *ctrl, int cnt) {
  apply_impulses(ctrl, cnt, SortedLinkRecord{g, keys, ms});
}

// Lists up to kLinkSortLds links whose keys fit 32 bits (grids below 2^32 / 27
// cells) are sorted in LDS and applied by one workgroup.  A longer list, or a
// larger grid, sets kErrLinksUnsorted and leaves the sphere untouched; the
// host then sorts it with rocprim and runs k_apply_links.
constexpr int kLinkSortLds = 8192;
constexpr int kRadixBits = 4, kRadixDigits = 1 << kRadixBits;
constexpr int kSortWaves = kApplyChunk / 64;
constexpr int kSortRounds = kLinkSortLds / kApplyChunk;  // list elements per thread
constexpr int kSortSlots = kSortRounds * kSortWaves;     // (round, wave) pairs
static_assert(kRadixDigits * kSortSlots == 2 * kApplyChunk, "two counts per thread in the scan");

struct LinkSortLds {
  uint32_t key[2][kLinkSortLds];
  uint16_t pos[2][kLinkSortLds];       // list position of the key (its momentum)
  int cnt[kRadixDigits][kSortSlots];   // digit-major, then (round, wave)
  int wsum[kSortWaves];
};

// Exclusive prefix sums, in place, of the 2 * kApplyChunk counts c (two
// consecutive counts per thread).
__device__ __forceinline__ void block_scan_pairs(int *c, int *wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int a = c[2 * tid], b = c[2 * tid + 1];
  int s = a + b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(s, o, 64);
    if (lane >= o) s += v;
  }
  if (lane == 63) wsum[wave] = s;
  __syncthreads();
  if (wave == 0) {
    int t = lane < kSortWaves ? wsum[lane] : 0;
#pragma unroll
    for (int o = 1; o < kSortWaves; o <<= 1) {
      const int v = __shfl_up(t, o, 64);
      if (lane >= o) t += v;
    }
    if (lane < kSortWaves) wsum[lane] = t;
  }
  __syncthreads();
  const int before = (wave > 0 ? wsum[wave - 1] : 0) + s - (a + b);
  c[2 * tid] = before;
  c[2 * tid + 1] = before + a;
  __syncthreads();
}

// Least-significant-digit radix sort of the first cnt keys of L.key[0] (and
// their positions) over their low kb bits, four bits per pass.  A pass is
// stable: element e = r * kApplyChunk + tid of round r goes after every
// element of a smaller digit and after the elements of its digit in earlier
// rounds, waves and lanes (ranks within a wave from ballots of the digit's
// bits).  Returns the buffer holding the sorted list.
__device__ __forceinline__ int radix_sort_lds(LinkSortLds &L, int cnt, int kb) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t below = lane ? (~(uint64_t)0 >> (64 - lane)) : 0;  // lanes under this one
  int src = 0;
  for (int shift = 0; shift < kb; shift += kRadixBits) {
    int *cnt_flat = &L.cnt[0][0];
    for (int i = tid; i < kRadixDigits * kSortSlots; i += kApplyChunk) cnt_flat[i] = 0;
    __syncthreads();
    uint32_t dig[kSortRounds];
    int rank[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; r++) {
      dig[r] = 0;
      rank[r] = 0;
      if (r * kApplyChunk >= cnt) continue;  // the same for every thread
      const int e = r * kApplyChunk + tid;
      const bool v = e < cnt;
      const uint32_t d = v ? (L.key[src][e] >> shift) & (kRadixDigits - 1) : 0u;
      uint64_t m = __ballot(v);
#pragma unroll
      for (int bit = 0; bit < kRadixBits; bit++) {
        const bool on = (d >> bit) & 1u;
        const uint64_t bb = __ballot(v && on);
        m &= on ? bb : ~bb;
      }
      const int rk = __popcll(m & below);
      if (v && rk == 0) L.cnt[d][r * kSortWaves + wave] = __popcll(m);
      dig[r] = d;
      rank[r] = rk;
    }
    __syncthreads();
    block_scan_pairs(cnt_flat, L.wsum);
#pragma unroll
    for (int r = 0; r < kSortRounds; r++) {
      const int e = r * kApplyChunk + tid;
      if (e >= cnt) continue;
      const int dst = L.cnt[dig[r]][r * kSortWaves + wave] + rank[r];
      L.key[src ^ 1][dst] = L.key[src][e];
      L.pos[src ^ 1][dst] = L.pos[src][e];
    }
    __syncthreads();
    src ^= 1;
  }
  return src;
}

struct LdsLinkRecord {
  Geom g;
  const uint32_t *key;
  const uint16_t *pos;
  const double *ms;
  __device__ void operator()(int j, int &x, int &y, int &z, double &px, double &py,
                             double &pz) const {
    link_impulse(g, (int64_t)key[j], ms[pos[j]], x, y, z, px, py, pz);
  }
};

__global__ __launch_bounds__(kApplyChunk) void k_sort_apply_links(Geom g,
                                                                  const int64_t *__restrict__ keys,
                                                                  const double *__restrict__ ms,
                                                                  Ctrl *ctrl, int cap) {
  __shared__ LinkSortLds L;
  const int tid = threadIdx.x;
  const int cnt = ctrl->n_links < cap ? ctrl->n_links : cap;
  const uint64_t max_key = (uint64_t)g.n * Q27 - 1;
  const int kb = max_key ? 64 - __clzll((long long)max_key) : 1;  // bits of the largest key
  if (cnt > kLinkSortLds || kb > 32) {
    if (tid == 0) atomicOr(&ctrl->err, kErrLinksUnsorted);
    return;
  }
  for (int t = tid; t < cnt; t += kApplyChunk) {
    L.key[0][t] = (uint32_t)keys[t];
    L.pos[0][t] = (uint16_t)t;
  }
  __syncthreads();
  const int b = radix_sort_lds(L, cnt, kb);
  apply_impulses(ctrl, cnt, LdsLinkRecord{g, L.key[b], L.pos[b], ms});
}

// absorbDeadCells (src/solid.js:87-101): the impulse rho_phase(phi) u of every
// dead cell of the last retag (src/solid.js:94-97), in the list's index order.
struct DeadRecord {
  Geom g;
  const int32_t *dead;
  double rhoG, dRho;
  const float *phi, *ux, *uy, *uz;
  __device__ void operator()(int j, int &x, int &y, int &z, double &px, double &py,
                             double &pz) const {
    const int32_t i = dead[j];
    const double rhoPhase = rhoG + (double)phi[i] * dRho;
    x = (int)(i % g.nx);
    y = (int)((i / g.nx) % g.ny);
    z = (int)(i / g.nxny);
    px = rhoPhase * (double)ux[i];
    py = rhoPhase * (double)uy[i];
    pz = rhoPhase * (double)uz[i];
  }
};

__global__ __launch_bounds__(kApplyChunk) void k_absorb_dead(Geom g, const int32_t *__restrict__ dead,
                                                             Ctrl *ctrl, int cap, double rhoG,
                                                             double dRho,
                                                             const float *__restrict__ phi,
                                                             const float *__restrict__ ux,
                                                             const float *__restrict__ uy,
                                                             const float *__restrict__ uz) {
  const int cnt = ctrl->n_dead < cap ? ctrl->n_dead : cap;
  apply_impulses(ctrl, cnt, DeadRecord{g, dead, rhoG, dRho, phi, ux, uy, uz});
}

// packVolume (src/sim.js:586-600) for one cell: R = (clamp(phi, 0, 1) * 255) | 0,
// G = 255 on solid cells; the two bytes as one little-endian 16-bit value.
__device__ __forceinline__ uint16_t pack_rg8(float phi, uint8_t tag) {
  double p = phi;
  if (p < 0) p = 0;
  else if (p > 1) p = 1;
  return (uint16_t)((uint8_t)(int32_t)(p * 255) | (tag == TAG_SOLID ? 0xff00u : 0u));
}

__global__ __launch_bounds__(256) void k_pack(int64_t n, const float *__restrict__ phi,
                                              const uint8_t *__restrict__ tag,
                                              uint8_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint16_t v = pack_rg8(phi[i], tag[i]);
  out[2 * i] = (uint8_t)v;
  out[2 * i + 1] = (uint8_t)(v >> 8);
}

// The frames of many scenes packed by one launch: grid row j packs the cells
// of job j into out + off[j] (an even byte offset), its blocks striding over
// them.
constexpr int kPackJobs = 32;
struct PackJobs {
  const float *phi[kPackJobs];
  const uint8_t *tag[kPackJobs];
  int64_t n[kPackJobs];
  int64_t off[kPackJobs];
};
__global__ __launch_bounds__(256) void k_pack_jobs(const PackJobs jobs, uint8_t *__restrict__ out) {
  const int j = blockIdx.y;
  const int64_t n = jobs.n[j];
  const float *__restrict__ phi = jobs.phi[j];
  const uint8_t *__restrict__ tag = jobs.tag[j];
  uint16_t *__restrict__ o = reinterpret_cast<uint16_t *>(out + jobs.off[j]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    o[i] = pack_rg8(phi[i], tag[i]);
}

// Layout transposes between the reference's array-of-structures (q values per
// cell) and the device's row blocks of float32 populations.
__global__ __launch_bounds__(256) void k_aos_to_rows(Geom g, int q, const float *__restrict__ aos,
                                                     float *__restrict__ rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int64_t pe = pop_base(g, i);
  for (int k = 0; k < q; k++) rows[pe + k * g.pb] = aos[i * q + k];
}
__global__ __launch_bounds__(256) void k_rows_to_aos(Geom g, int q, const float *__restrict__ rows,
                                                     float *__restrict__ aos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const int64_t pe = pop_base(g, i);
  for (int k = 0; k < q; k++) aos[i * q + k] = rows[pe + k * g.pb];
}

}  // namespace lbm
