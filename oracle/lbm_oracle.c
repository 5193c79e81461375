/*
 * CPU oracle for the coupled LBM step -- TEST INFRASTRUCTURE ONLY (see lbm_oracle.h).
 *
 * Each function cites the reference function it restates.  Expressions keep the
 * reference's evaluation order (JavaScript numbers are IEEE doubles, evaluated
 * left to right); stores round to ostore_t exactly where the reference stores
 * into a Float32Array.  Integer lattice components enter products as exact
 * doubles, as the reference's Int8Array elements do.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Built with -fopenmp (liboracle32_omp.so, a CPU timing baseline) the per-cell
 * passes run their z-planes in parallel: each cell writes only its own
 * elements, and the one order-dependent sum (momentum exchange into the body)
 * is applied after the pass in cell order, so every thread count gives the
 * sequential result bit for bit. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

enum { Q27 = 27, Q7 = 7, TAG_FLUID = 0, TAG_SOLID = 1, TAG_WALL = 2 };

/* lattices.js:34-49 -- direction order, opposite pairs adjacent. */
static const int8_t EX27[Q27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                                 1, -1, 1, -1, 1, -1, -1, 1};
static const int8_t EY27[Q27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1,
                                 1, -1, 1, -1, -1, 1, 1, -1};
static const int8_t EZ27[Q27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1,
                                 1, -1, -1, 1, 1, -1, 1, -1};
/* lattices.js:79-84 */
static const int8_t EX7[Q7] = {0, 1, -1, 0, 0, 0, 0};
static const int8_t EY7[Q7] = {0, 0, 0, 1, -1, 0, 0};
static const int8_t EZ7[Q7] = {0, 0, 0, 0, 0, 1, -1};

/* lattices.js:22,52-57: weights are stored in a Float32Array, so each is the
 * double 8/27 (etc.) rounded to float32. */
static float W27[Q27];
static float W7[Q7];
static int8_t OPP27[Q27], OPP7[Q7];
static int tables_ready = 0;

static const double INV_CS2_27 = 3.0;  /* lattices.js:96 */
static const double INV_2CS4_27 = 4.5; /* lattices.js:97 */
static const double CS2_7 = 1.0 / 4;   /* lattices.js:98 */
static const double INV_CS2_7 = 4.0;   /* lattices.js:99 */

static void build_tables(void) {
  if (tables_ready) return;
  for (int k = 0; k < Q27; k++) {
    double w;
    if (k == 0) w = 8.0 / 27;
    else if (k <= 6) w = 2.0 / 27;
    else if (k <= 18) w = 1.0 / 54;
    else w = 1.0 / 216;
    W27[k] = (float)w;
    /* lattices.js:66-69: OPP = k odd ? k+1 : k-1 (0 is its own opposite). */
    OPP27[k] = (int8_t)(k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1));
  }
  const double w7v[Q7] = {1.0 / 4, 1.0 / 8, 1.0 / 8, 1.0 / 8, 1.0 / 8, 1.0 / 8, 1.0 / 8};
  for (int k = 0; k < Q7; k++) {
    W7[k] = (float)w7v[k];
    OPP7[k] = (int8_t)(k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1)); /* lattices.js:90-91 */
  }
  tables_ready = 1;
}

int oracle_tables(int8_t *ex27, int8_t *ey27, int8_t *ez27, float *w27, int8_t *opp27,
                  int8_t *ex7, int8_t *ey7, int8_t *ez7, float *w7, int8_t *opp7) {
  build_tables();
  memcpy(ex27, EX27, Q27);
  memcpy(ey27, EY27, Q27);
  memcpy(ez27, EZ27, Q27);
  memcpy(w27, W27, sizeof W27);
  memcpy(opp27, OPP27, Q27);
  memcpy(ex7, EX7, Q7);
  memcpy(ey7, EY7, Q7);
  memcpy(ez7, EZ7, Q7);
  memcpy(w7, W7, sizeof W7);
  memcpy(opp7, OPP7, Q7);
  return 0;
}

int oracle_storage_bytes(void) { return (int)sizeof(ostore_t); }

static ostore_t *zalloc(size_t count) { return (ostore_t *)calloc(count, sizeof(ostore_t)); }

/* sim.js:30-96 */
osim *oracle_sim_create(int nx, int ny, int nz) {
  build_tables();
  osim *s = (osim *)calloc(1, sizeof(osim));
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  const int n = nx * ny * nz;
  s->n = n;
  s->f = zalloc((size_t)n * Q27);
  s->f2 = zalloc((size_t)n * Q27);
  s->h = zalloc((size_t)n * Q7);
  s->h2 = zalloc((size_t)n * Q7);
  s->rho = zalloc(n);
  s->ux = zalloc(n);
  s->uy = zalloc(n);
  s->uz = zalloc(n);
  s->phi = zalloc(n);
  s->phiPrev = zalloc(n);
  s->tag = (uint8_t *)calloc(n, 1);
  s->tagPrev = (uint8_t *)calloc(n, 1);
  s->usx = zalloc(n);
  s->usy = zalloc(n);
  s->usz = zalloc(n);
  s->gx = zalloc(n);
  s->gy = zalloc(n);
  s->gz = zalloc(n);
  s->lap = zalloc(n);
  s->bodyFx = zalloc(n);
  s->bodyFy = zalloc(n);
  s->bodyFz = zalloc(n);
  const int nxny = nx * ny;
  for (int k = 0; k < Q27; k++) s->neighOffset27[k] = -(EZ27[k] * nxny + EY27[k] * nx + EX27[k]);
  for (int k = 0; k < Q7; k++) s->neighOffset7[k] = -(EZ7[k] * nxny + EY7[k] * nx + EX7[k]);
  /* sim.js:88-95 defaults */
  s->tau = 0.6;
  s->tauPhi = 0.7;
  s->sigma = 0.012;
  s->gravity = 0.0010;
  s->rhoL = 1.0;
  s->rhoG = 0.05;
  s->W = 4.0;
  s->mobility = 0.02;
  return s;
}

void oracle_sim_destroy(osim *s) {
  if (!s) return;
  ostore_t *arrs[] = {s->f,   s->f2,  s->h,   s->h2,  s->rho, s->ux,     s->uy,     s->uz,
                      s->phi, s->phiPrev, s->usx, s->usy, s->usz, s->gx, s->gy, s->gz,
                      s->lap, s->bodyFx, s->bodyFy, s->bodyFz};
  for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; i++) free(arrs[i]);
  free(s->tag);
  free(s->tagPrev);
  free(s);
}

static inline int idx(const osim *s, int x, int y, int z) { return ((z * s->ny) + y) * s->nx + x; }

/* sim.js:112-134 body, with the per-column phi supplied by the caller. */
static void init_columns(osim *s, const double *yw_xz, double waterlineY) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double W = s->W;
  for (int z = 0; z < nz; z++) {
    for (int y = 0; y < ny; y++) {
      for (int x = 0; x < nx; x++) {
        const double wl = yw_xz ? yw_xz[z * nx + x] : waterlineY;
        const double phiCol = 0.5 * (1 + tanh(2 * (y - wl) / W));
        const int i = idx(s, x, y, z);
        s->phi[i] = (ostore_t)phiCol;
        s->phiPrev[i] = (ostore_t)phiCol;
        s->rho[i] = (ostore_t)1.0;
        s->ux[i] = s->uy[i] = s->uz[i] = 0;
        const int onWall = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
        s->tag[i] = onWall ? TAG_WALL : TAG_FLUID;
        s->tagPrev[i] = s->tag[i];
        for (int k = 0; k < Q27; k++) s->f[(size_t)i * Q27 + k] = (ostore_t)W27[k];
        for (int k = 0; k < Q7; k++) s->h[(size_t)i * Q7 + k] = (ostore_t)((double)W7[k] * phiCol);
      }
    }
  }
  oracle_seed_hydrostatic(s);
}

/* sim.js:112-134 initFlat */
void oracle_init_flat(osim *s, double waterlineY) { init_columns(s, NULL, waterlineY); }

/* initFlat generalised to one waterline per (x,z) column (yw_xz[z*nx+x]); used
 * for perturbed interfaces.  With a constant map it equals oracle_init_flat. */
void oracle_init_waterline_map(osim *s, const double *yw_xz) { init_columns(s, yw_xz, 0.0); }

/* sim.js:144-173 seedHydrostatic */
void oracle_seed_hydrostatic(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double dRho = s->rhoL - s->rhoG;
  const double gOverCs2 = s->gravity * INV_CS2_27;
  const double uLatY = -0.5 * s->gravity;
  const double uLat2 = uLatY * uLatY;
  for (int z = 0; z < nz; z++) {
    for (int x = 0; x < nx; x++) {
      double pCs2 = 1.0;
      for (int y = 0; y < ny; y++) {
        const int i = idx(s, x, y, z);
        s->rho[i] = (ostore_t)pCs2;
        const size_t i27 = (size_t)i * Q27;
        const double rhoPhase = s->rhoG + (double)s->phi[i] * dRho;
        for (int k = 0; k < Q27; k++) {
          const double eu = (double)EY27[k] * uLatY;
          const double Qi = eu * (INV_CS2_27 + INV_2CS4_27 * eu) - 1.5 * uLat2;
          s->f[i27 + k] = (ostore_t)((double)W27[k] * (pCs2 + rhoPhase * Qi));
        }
        pCs2 += rhoPhase * gOverCs2;
      }
    }
  }
}

/* sim.js:176-199 computePhiGradients: central differences, indices clamped. */
void oracle_compute_phi_gradients(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const ostore_t *phi = s->phi;
  OMP_FOR
  for (int z = 0; z < nz; z++) {
    const int zm = z > 0 ? z - 1 : z, zp = z < nz - 1 ? z + 1 : z;
    for (int y = 0; y < ny; y++) {
      const int ym = y > 0 ? y - 1 : y, yp = y < ny - 1 ? y + 1 : y;
      for (int x = 0; x < nx; x++) {
        const int xm = x > 0 ? x - 1 : x, xp = x < nx - 1 ? x + 1 : x;
        const int i = ((z * ny) + y) * nx + x;
        const double c = phi[i];
        const double px = phi[((z * ny) + y) * nx + xp];
        const double mx = phi[((z * ny) + y) * nx + xm];
        const double py = phi[((z * ny) + yp) * nx + x];
        const double my = phi[((z * ny) + ym) * nx + x];
        const double pz = phi[((zp * ny) + y) * nx + x];
        const double mz = phi[((zm * ny) + y) * nx + x];
        s->gx[i] = (ostore_t)(0.5 * (px - mx));
        s->gy[i] = (ostore_t)(0.5 * (py - my));
        s->gz[i] = (ostore_t)(0.5 * (pz - mz));
        s->lap[i] = (ostore_t)(px + mx + py + my + pz + mz - 6 * c);
      }
    }
  }
}

/* sim.js:214-244 retagCells (also snapshots tag and phi). */
void oracle_retag_cells(osim *s, const osphere *solid) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  memcpy(s->tagPrev, s->tag, (size_t)s->n);
  memcpy(s->phiPrev, s->phi, (size_t)s->n * sizeof(ostore_t));
  const int nxny = nx * ny;
  OMP_FOR
  for (int z = 0; z < nz; z++) {
    double tmp[3] = {0, 0, 0};
    const int zOnWall = z == 0 || z == nz - 1;
    const int zBase = z * nxny;
    for (int y = 0; y < ny; y++) {
      const int yOnWall = y == 0 || y == ny - 1;
      const int yzBase = zBase + y * nx;
      for (int x = 0; x < nx; x++) {
        const int i = yzBase + x;
        if (zOnWall || yOnWall || x == 0 || x == nx - 1) {
          s->tag[i] = TAG_WALL;
          s->usx[i] = s->usy[i] = s->usz[i] = 0;
          continue;
        }
        if (solid && oracle_sphere_contains(solid, x, y, z)) {
          s->tag[i] = TAG_SOLID;
          oracle_sphere_surface_velocity(solid, x, y, z, tmp);
          s->usx[i] = (ostore_t)tmp[0];
          s->usy[i] = (ostore_t)tmp[1];
          s->usz[i] = (ostore_t)tmp[2];
        } else {
          s->tag[i] = TAG_FLUID;
          s->usx[i] = s->usy[i] = s->usz[i] = 0;
        }
      }
    }
  }
}

/* sim.js:257-310 handleFreshCells: in-place, z-y-x order, so a fresh cell sees
 * the already-reseeded values of earlier fresh neighbours. */
void oracle_handle_fresh_cells(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  for (int z = 1; z < nz - 1; z++) {
    for (int y = 1; y < ny - 1; y++) {
      for (int x = 1; x < nx - 1; x++) {
        const int i = ((z * ny) + y) * nx + x;
        if (s->tag[i] != TAG_FLUID || s->tagPrev[i] != TAG_SOLID) continue;
        double psum = 0, rsum = 0, vxs = 0, vys = 0, vzs = 0;
        int pcnt = 0;
        for (int dz = -1; dz <= 1; dz++) {
          for (int dy = -1; dy <= 1; dy++) {
            for (int dx = -1; dx <= 1; dx++) {
              if (!dx && !dy && !dz) continue;
              const int j = (((z + dz) * ny) + (y + dy)) * nx + (x + dx);
              if (s->tag[j] == TAG_FLUID) {
                psum += s->phi[j];
                rsum += s->rho[j];
                vxs += s->ux[j];
                vys += s->uy[j];
                vzs += s->uz[j];
                pcnt++;
              }
            }
          }
        }
        const double p = pcnt > 0 ? psum / pcnt : (double)s->phiPrev[i];
        const double r = pcnt > 0 ? rsum / pcnt : 1.0;
        const double vx = pcnt > 0 ? vxs / pcnt : 0;
        const double vy = pcnt > 0 ? vys / pcnt : 0;
        const double vz = pcnt > 0 ? vzs / pcnt : 0;
        s->phi[i] = (ostore_t)p;
        s->phiPrev[i] = (ostore_t)p;
        s->rho[i] = (ostore_t)r;
        s->ux[i] = (ostore_t)vx;
        s->uy[i] = (ostore_t)vy;
        s->uz[i] = (ostore_t)vz;
        const double rhoPhase = s->rhoG + p * (s->rhoL - s->rhoG);
        const double u2 = vx * vx + vy * vy + vz * vz;
        for (int k = 0; k < Q27; k++) {
          const double eu = EX27[k] * vx + EY27[k] * vy + EZ27[k] * vz;
          const double Qi = eu * (INV_CS2_27 + INV_2CS4_27 * eu) - 1.5 * u2;
          s->f[(size_t)i * Q27 + k] = (ostore_t)((double)W27[k] * (r + rhoPhase * Qi));
        }
        for (int k = 0; k < Q7; k++) {
          const double eu = EX7[k] * vx + EY7[k] * vy + EZ7[k] * vz;
          s->h[(size_t)i * Q7 + k] = (ostore_t)((double)W7[k] * p * (1 + INV_CS2_7 * eu));
        }
      }
    }
  }
}

/* sim.js:324-364 collidePhase: BGK + anti-diffusion source. */
void oracle_collide_phase(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double tauPhi = s->tauPhi, W = s->W;
  const double invTau = 1 / tauPhi;
  const int nxny = nx * ny;
  const double Mphi = CS2_7 * (tauPhi - 0.5);
  const double guoHalf = 1.0 - 0.5 * invTau;
  const double sharpC = 4.0 * Mphi * guoHalf / W;
  const double sharpFac = INV_CS2_7;
  OMP_FOR
  for (int z = 1; z < nz - 1; z++) {
    const int zBase = z * nxny;
    for (int y = 1; y < ny - 1; y++) {
      const int yzBase = zBase + y * nx;
      for (int x = 1; x < nx - 1; x++) {
        const int i = yzBase + x;
        if (s->tag[i] != TAG_FLUID) continue;
        const size_t i7 = (size_t)i * Q7;
        const double p = s->phi[i];
        const double uxi = s->ux[i], uyi = s->uy[i], uzi = s->uz[i];
        const double ggx = s->gx[i], ggy = s->gy[i], ggz = s->gz[i];
        const double gmag = sqrt(ggx * ggx + ggy * ggy + ggz * ggz) + 1e-12;
        const double sharpen = sharpC * p * (1 - p);
        const double Sx = sharpen * sharpFac * ggx / gmag;
        const double Sy = sharpen * sharpFac * ggy / gmag;
        const double Sz = sharpen * sharpFac * ggz / gmag;
        for (int k = 0; k < Q7; k++) {
          const double ekx = EX7[k], eky = EY7[k], ekz = EZ7[k];
          const double wk = W7[k];
          const double eu = ekx * uxi + eky * uyi + ekz * uzi;
          const double heq = wk * p * (1 + INV_CS2_7 * eu);
          const double Fk = wk * (ekx * Sx + eky * Sy + ekz * Sz);
          const double hk = s->h[i7 + k];
          s->h2[i7 + k] = (ostore_t)(hk - invTau * (hk - heq) + Fk);
        }
      }
    }
  }
}

/* sim.js:376-397 streamPhase: pull; non-fluid source bounces (no wall term). */
void oracle_stream_phase(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const int nxny = nx * ny;
  OMP_FOR
  for (int z = 1; z < nz - 1; z++) {
    const int zBase = z * nxny;
    for (int y = 1; y < ny - 1; y++) {
      const int yzBase = zBase + y * nx;
      for (int x = 1; x < nx - 1; x++) {
        const int i = yzBase + x;
        if (s->tag[i] != TAG_FLUID) continue;
        const size_t i7 = (size_t)i * Q7;
        for (int k = 0; k < Q7; k++) {
          const int sIdx = i + s->neighOffset7[k];
          s->h[i7 + k] = (s->tag[sIdx] == TAG_FLUID) ? s->h2[(size_t)sIdx * Q7 + k]
                                                      : s->h2[i7 + OPP7[k]];
        }
      }
    }
  }
}

/* sim.js:406-464 collideHydro: HCZ equilibrium + Guo source. */
void oracle_collide_hydro(osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double tau = s->tau, sigma = s->sigma, gravity = s->gravity;
  const double rhoL = s->rhoL, rhoG = s->rhoG, W = s->W;
  const double invTau = 1 / tau;
  const double beta = 12 * sigma / W;
  const double kappa = 1.5 * sigma * W;
  const double dRho = rhoL - rhoG;
  const double A = (1 - 0.5 * invTau) * INV_CS2_27;
  const int nxny = nx * ny;
  OMP_FOR
  for (int z = 1; z < nz - 1; z++) {
    const int zBase = z * nxny;
    for (int y = 1; y < ny - 1; y++) {
      const int yzBase = zBase + y * nx;
      for (int x = 1; x < nx - 1; x++) {
        const int i = yzBase + x;
        if (s->tag[i] != TAG_FLUID) continue;
        const size_t i27 = (size_t)i * Q27;
        const double sumF = s->rho[i];
        const double p = s->phi[i];
        const double rhoPhase = rhoG + p * dRho;
        const double mu = 4 * beta * p * (p - 1) * (2 * p - 1) - kappa * (double)s->lap[i];
        const double fx = (double)s->bodyFx[i] + mu * (double)s->gx[i];
        const double fy = (double)s->bodyFy[i] + mu * (double)s->gy[i] + rhoPhase * gravity;
        const double fz = (double)s->bodyFz[i] + mu * (double)s->gz[i];
        const double uxe = s->ux[i];
        const double uye = s->uy[i];
        const double uze = s->uz[i];
        const double u2 = uxe * uxe + uye * uye + uze * uze;
        const double negU2half = -1.5 * u2;
        const double uF = uxe * fx + uye * fy + uze * fz;
        for (int k = 0; k < Q27; k++) {
          const double ekx = EX27[k], eky = EY27[k], ekz = EZ27[k];
          const double wk = W27[k];
          const double eu = ekx * uxe + eky * uye + ekz * uze;
          const double Qi = negU2half + eu * (INV_CS2_27 + INV_2CS4_27 * eu);
          const double feq = wk * (sumF + rhoPhase * Qi);
          const double ekF = ekx * fx + eky * fy + ekz * fz;
          const double Sk = A * wk * (ekF * (1 + INV_CS2_27 * eu) - uF);
          const double fk = s->f[i27 + k];
          s->f2[i27 + k] = (ostore_t)(fk - invTau * (fk - feq) + Sk);
        }
      }
    }
  }
}

/* Momentum-exchange impulses of one z-plane, in cell order. */
typedef struct {
  int x, y, z;
  double px, py, pz;
} oimpulse;
typedef struct {
  oimpulse *v;
  size_t n, cap;
} oimplist;
static void imp_push(oimplist *l, int x, int y, int z, double px, double py, double pz) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 256;
    l->v = (oimpulse *)realloc(l->v, l->cap * sizeof *l->v);
    if (!l->v) {
      fprintf(stderr, "oracle: out of memory\n");
      abort();
    }
  }
  oimpulse *e = &l->v[l->n++];
  e->x = x;
  e->y = y;
  e->z = z;
  e->px = px;
  e->py = py;
  e->pz = pz;
}

/* sim.js:484-582 streamHydroAndMacro: pull stream, halfway bounce-back with the
 * moving-wall term, momentum exchange into the body, then rho / phi / u.  The
 * impulses are recorded per plane and applied after the cell loop in the
 * reference's cell order (they only add to the body's force and torque, which
 * the loop does not read). */
void oracle_stream_hydro_and_macro(osim *s, osphere *solid) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double sigma = s->sigma, gravity = s->gravity;
  const double rhoL = s->rhoL, rhoG = s->rhoG, W = s->W;
  const int nxny = nx * ny;
  const double beta = 12 * sigma / W;
  const double kappa = 1.5 * sigma * W;
  const double dRho = rhoL - rhoG;
  oimplist *imp = (oimplist *)calloc((size_t)nz, sizeof *imp);
  if (!imp) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  OMP_FOR
  for (int z = 1; z < nz - 1; z++) {
    const int zBase = z * nxny;
    for (int y = 1; y < ny - 1; y++) {
      const int yzBase = zBase + y * nx;
      for (int x = 1; x < nx - 1; x++) {
        const int i = yzBase + x;
        if (s->tag[i] != TAG_FLUID) continue;
        const size_t i27 = (size_t)i * Q27;
        const double rhoPhaseOld = rhoG + (double)s->phi[i] * dRho;
        double pSumF = 0, mx = 0, my = 0, mz = 0;
        for (int k = 0; k < Q27; k++) {
          const int sIdx = i + s->neighOffset27[k];
          const int sTag = s->tag[sIdx];
          const double ekx = EX27[k], eky = EY27[k], ekz = EZ27[k];
          double fk;
          if (sTag == TAG_FLUID) {
            fk = s->f2[(size_t)sIdx * Q27 + k];
          } else {
            const double fopp = s->f2[i27 + OPP27[k]];
            double corr = 0;
            if (sTag == TAG_SOLID) {
              const double eu = ekx * (double)s->usx[sIdx] + eky * (double)s->usy[sIdx] +
                                ekz * (double)s->usz[sIdx];
              corr = 2 * (double)W27[k] * rhoPhaseOld * eu * INV_CS2_27;
            }
            fk = fopp + corr;
            if (sTag == TAG_SOLID && solid) {
              const double m = fopp + fk;
              const int sx = x - (int)ekx, sy = y - (int)eky, sz = z - (int)ekz;
              imp_push(&imp[z], sx, sy, sz, -m * ekx, -m * eky, -m * ekz);
            }
          }
          s->f[i27 + k] = (ostore_t)fk;
          pSumF += fk;
          mx += ekx * fk;
          my += eky * fk;
          mz += ekz * fk;
        }
        s->rho[i] = (ostore_t)(pSumF > 0.01 ? pSumF : 0.01);
        double phiSum = 0;
        for (int k = 0; k < Q7; k++) phiSum += s->h[(size_t)i * Q7 + k];
        const double p = phiSum < 0 ? 0 : phiSum > 1 ? 1 : phiSum;
        s->phi[i] = (ostore_t)p;
        const double rhoPhaseNew = rhoG + p * dRho;
        const double mu = 4 * beta * p * (p - 1) * (2 * p - 1) - kappa * (double)s->lap[i];
        const double fx = (double)s->bodyFx[i] + mu * (double)s->gx[i];
        const double fy = (double)s->bodyFy[i] + mu * (double)s->gy[i] + rhoPhaseNew * gravity;
        const double fz = (double)s->bodyFz[i] + mu * (double)s->gz[i];
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
        s->ux[i] = (ostore_t)uxn;
        s->uy[i] = (ostore_t)uyn;
        s->uz[i] = (ostore_t)uzn;
      }
    }
  }
  for (int z = 0; z < nz; z++) {
    for (size_t j = 0; j < imp[z].n; j++) {
      const oimpulse *e = &imp[z].v[j];
      oracle_sphere_apply_impulse_at_cell(solid, e->x, e->y, e->z, e->px, e->py, e->pz);
    }
    free(imp[z].v);
  }
  free(imp);
  memset(s->bodyFx, 0, (size_t)s->n * sizeof(ostore_t));
  memset(s->bodyFy, 0, (size_t)s->n * sizeof(ostore_t));
  memset(s->bodyFz, 0, (size_t)s->n * sizeof(ostore_t));
}

/* sim.js:202-210 step: the fixed pass order. */
void oracle_sim_step(osim *s, osphere *solid) {
  oracle_retag_cells(s, solid);
  oracle_handle_fresh_cells(s);
  oracle_compute_phi_gradients(s);
  oracle_collide_phase(s);
  oracle_stream_phase(s);
  oracle_collide_hydro(s);
  oracle_stream_hydro_and_macro(s, solid);
}

/* sim.js:586-600 packVolume: R = (clamp(phi)*255)|0, G = 255 on SOLID. */
void oracle_pack_volume(const osim *s, uint8_t *out) {
  size_t o = 0;
  for (int z = 0; z < s->nz; z++)
    for (int y = 0; y < s->ny; y++)
      for (int x = 0; x < s->nx; x++) {
        const int i = ((z * s->ny) + y) * s->nx + x;
        double p = s->phi[i];
        if (p < 0) p = 0;
        else if (p > 1) p = 1;
        out[o++] = (uint8_t)(int32_t)(p * 255);
        out[o++] = s->tag[i] == TAG_SOLID ? 255 : 0;
      }
}

/* ------------------------------------------------------------------ solid.js */

/* solid.js:9-37 constructor with explicit options (the caller resolves the
 * reference's defaults: cx=nx/2, cy=0.3ny, cz=nz/2, r=0.1 min(n), density 1.4). */
void oracle_sphere_init(osphere *b, int nx, int ny, int nz, double cx, double cy, double cz,
                        double r, double density) {
  memset(b, 0, sizeof *b);
  b->nx = nx;
  b->ny = ny;
  b->nz = nz;
  b->cx = cx;
  b->cy = cy;
  b->cz = cz;
  b->r = r;
  b->density = density;
  oracle_sphere_recompute_mass(b);
}

/* solid.js:39-44 */
void oracle_sphere_recompute_mass(osphere *b) {
  const double V = (4.0 / 3) * M_PI * b->r * b->r * b->r;
  b->volume = V;
  b->mass = b->density * V;
  b->I = (2.0 / 5) * b->mass * b->r * b->r;
}

/* solid.js:50-55 */
int oracle_sphere_contains(const osphere *b, int x, int y, int z) {
  const double dx = x + 0.5 - b->cx;
  const double dy = y + 0.5 - b->cy;
  const double dz = z + 0.5 - b->cz;
  return dx * dx + dy * dy + dz * dz <= b->r * b->r;
}

/* solid.js:57-66 */
void oracle_sphere_surface_velocity(const osphere *b, int x, int y, int z, double *out) {
  const double rx = x + 0.5 - b->cx;
  const double ry = y + 0.5 - b->cy;
  const double rz = z + 0.5 - b->cz;
  out[0] = b->vx + (b->wy * rz - b->wz * ry);
  out[1] = b->vy + (b->wz * rx - b->wx * rz);
  out[2] = b->vz + (b->wx * ry - b->wy * rx);
}

/* solid.js:70-80 */
void oracle_sphere_apply_impulse_at_cell(osphere *b, int x, int y, int z, double px, double py,
                                         double pz) {
  b->Fx += px;
  b->Fy += py;
  b->Fz += pz;
  const double rx = x + 0.5 - b->cx;
  const double ry = y + 0.5 - b->cy;
  const double rz = z + 0.5 - b->cz;
  b->Tx += ry * pz - rz * py;
  b->Ty += rz * px - rx * pz;
  b->Tz += rx * py - ry * px;
}

/* solid.js:87-101 */
void oracle_sphere_absorb_dead_cells(osphere *b, const osim *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double dRho = s->rhoL - s->rhoG;
  for (int z = 1; z < nz - 1; z++)
    for (int y = 1; y < ny - 1; y++)
      for (int x = 1; x < nx - 1; x++) {
        const int i = ((z * ny) + y) * nx + x;
        if (s->tag[i] != 1 || s->tagPrev[i] != 0) continue;
        const double rhoPhase = s->rhoG + (double)s->phi[i] * dRho;
        oracle_sphere_apply_impulse_at_cell(b, x, y, z, rhoPhase * (double)s->ux[i],
                                            rhoPhase * (double)s->uy[i], rhoPhase * (double)s->uz[i]);
      }
}

/* Math.max(0, v) and Math.min(0, v) including NaN and signed-zero rules. */
static inline double js_max0(double v) { return v != v ? v : (v > 0 ? v : 0.0); }
static inline double js_min0(double v) {
  if (v != v) return v;
  if (v < 0) return v;
  return (v == 0 && signbit(v)) ? -0.0 : 0.0;
}

/* solid.js:104-210 step: gravity, drag spring, capped linear acceleration,
 * angular update, speed cap, damping, snap, move, clamp, reset. */
void oracle_sphere_step(osphere *b, double dt, double gravity) {
  const double bodyG = gravity;
  b->Fy += b->mass * bodyG;
  if (b->dragging != 0) {
    const double k = 0.004 * b->mass;
    const double cD = 2.0 * sqrt(k * b->mass);
    b->Fx += k * (b->dragTargetX - b->cx) - cD * b->vx;
    b->Fy += k * (b->dragTargetY - b->cy) - cD * b->vy;
    b->Fz += k * (b->dragTargetZ - b->cz) - cD * b->vz;
  }
  /* Math.abs(gravity || 0.001): 0 and NaN are falsy. */
  const double gsel = (gravity != 0 && gravity == gravity) ? gravity : 0.001;
  const double accCap = 3.5 * fabs(gsel);
  {
    const double ax = b->Fx / b->mass;
    const double ay = b->Fy / b->mass;
    const double az = b->Fz / b->mass;
    const double a2 = ax * ax + ay * ay + az * az;
    const double c2 = accCap * accCap;
    const double sc = a2 > c2 ? accCap / sqrt(a2) : 1.0;
    b->vx += dt * ax * sc;
    b->vy += dt * ay * sc;
    b->vz += dt * az * sc;
  }
  b->wx += dt * b->Tx / b->I;
  b->wy += dt * b->Ty / b->I;
  b->wz += dt * b->Tz / b->I;
  const double BODY_UMAX = 0.08;
  const double sp2 = b->vx * b->vx + b->vy * b->vy + b->vz * b->vz;
  if (sp2 > BODY_UMAX * BODY_UMAX) {
    const double sc = BODY_UMAX / sqrt(sp2);
    b->vx *= sc;
    b->vy *= sc;
    b->vz *= sc;
  }
  const double linDamp = 0.98, omegaDamp = 0.95;
  b->vx *= linDamp;
  b->vy *= linDamp;
  b->vz *= linDamp;
  b->wx *= omegaDamp;
  b->wy *= omegaDamp;
  b->wz *= omegaDamp;
  const double SNAP = 0.0001;
  if (fabs(b->vx) < SNAP) b->vx = 0;
  if (fabs(b->vy) < SNAP) b->vy = 0;
  if (fabs(b->vz) < SNAP) b->vz = 0;
  if (fabs(b->wx) < SNAP) b->wx = 0;
  if (fabs(b->wy) < SNAP) b->wy = 0;
  if (fabs(b->wz) < SNAP) b->wz = 0;
  b->cx += dt * b->vx;
  b->cy += dt * b->vy;
  b->cz += dt * b->vz;
  const double m = b->r + 1.5;
  if (b->cx < m) { b->cx = m; b->vx = js_max0(b->vx); }
  if (b->cx > b->nx - 1 - m) { b->cx = b->nx - 1 - m; b->vx = js_min0(b->vx); }
  if (b->cy < m) { b->cy = m; b->vy = js_max0(b->vy); }
  if (b->cy > b->ny - 1 - m) { b->cy = b->ny - 1 - m; b->vy = js_min0(b->vy); }
  if (b->cz < m) { b->cz = m; b->vz = js_max0(b->vz); }
  if (b->cz > b->nz - 1 - m) { b->cz = b->nz - 1 - m; b->vz = js_min0(b->vz); }
  b->Fx = b->Fy = b->Fz = 0;
  b->Tx = b->Ty = b->Tz = 0;
}

/* worker.js:96-100 / test/bench.mjs:22-26 driver order. */
void oracle_coupled_steps(osim *s, osphere *b, int nsteps) {
  for (int t = 0; t < nsteps; t++) {
    oracle_sim_step(s, b);
    if (b) {
      oracle_sphere_absorb_dead_cells(b, s);
      oracle_sphere_step(b, 1.0, s->gravity);
    }
  }
}
