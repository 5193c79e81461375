// Row march of the fused float32 step (one scene): the phase stream and the
// stream + collide kernel of a step in one launch.
//
// collide(t+1) of a cell needs phi(t+1) of its six face neighbours
// (src/sim.js:176-199 on the phi that streamPhase leaves, :376-397), and
// phi(t+1) of a cell is the clamped sum of the h pulled into it.  The split
// form computes it for every cell in a pass of its own (km_phase_stream),
// which reads every h a second time.  Here a workgroup owns one grid row y
// (up to 256 cells, one per lane) over a run of planes and marches in z: each
// iteration computes phi(t+1) of plane z + 1 for its own row and for rows
// y - 1 and y + 1, keeps the own row's pulled h in LDS, and then streams,
// updates and collides plane z with the whole phi(t+1) stencil taken from
// LDS.  Rows y +- 1 are computed redundantly: the workgroups that own them do
// the same loads at about the same time on the same XCD, so those loads hit
// L2.  Every h is fetched from HBM once per step and phi(t+1) is written once
// and never read back within the step.
//
// The arithmetic is that of km_phase_stream and stream_collide_body, in the
// same order, so the state after a step is the split form's bit for bit.
// Cells of the band box are left to k1_stream_collide_box (their tags, solid
// links and momentum), but their phi(t+1) is computed and stored here, as the
// phase pass did.
#pragma once

#include "kernels_fused.h"

namespace lbm {

constexpr int kMarchCells = 256;  // cells of a row a workgroup covers (one per lane)

// LDS of a marching workgroup of RPW rows.  Plane z of an own row's phi(t+1)
// sits in slot z & 3 of p, of a halo row and of the own rows' pulled h in slot
// z % 3: with one barrier per iteration, a wave that writes plane z + 2 while
// another still reads planes z - 1 .. z + 1 never meets it.  Halo rows: with
// one row y, hal[0] holds row y - 1 and hal[1] row y + 1; with two rows
// (y, y + 1), hal[0] holds row y - 1 and hal[1] row y + 2.
template <int RPW>
struct MarchLds {
  float p[RPW][4][kMarchCells];
  float hal[2][3][kMarchCells];
  float h[RPW][3][Q7][kMarchCells];
};

// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a full __syncthreads fence would wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// phi(t+1) of cell (x, y, z) as km_phase_stream computes it: the clamped sum of
// the pulled h of a fluid cell, phi(t) copied for walls and solid cells.  hv
// receives the pulled h and fluid whether the cell is fluid.
__device__ __forceinline__ float march_phase(const Geom &g, const int box[6],
                                             const uint8_t *__restrict__ tag,
                                             const float *__restrict__ h2,
                                             const float *__restrict__ phi, int x, int y, int z,
                                             float hv[Q7], bool &fluid) {
  const int64_t i = ((int64_t)z * g.ny + y) * g.nx + x;
  const uint32_t bo4 = (uint32_t)i * 4u;
  const bool near = in_box(box, x, y, z);
  fluid = interior(g, x, y, z) && (!near || ldb(tag, (uint32_t)i) == TAG_FLUID);
  if (!fluid) return ldb(phi, bo4);
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
  pull7<ST_F32>(g, row_addr(g, i, z * g.ny + y), h2, nonfluid, hv);
  double s = 0;
#pragma unroll
  for (int k = 0; k < Q7; k++) s += (double)hv[k];
  const double pp = s < 0 ? 0.0 : s > 1 ? 1.0 : s;
  return (float)pp;
}

// A workgroup of RPW * 256 lanes owns RPW consecutive rows (lanes 256 r + x:
// cell x of row r) over a segment of planes.  Workgroup b of a launch of
// 8 * cols8 * nseg: the rows come in units of RPW, and XCD b & 7 owns units
// [xcd U / 8, (xcd + 1) U / 8) of the U = ceil(ny / RPW) (the hardware deals
// consecutive workgroups to the eight XCDs in turn), so the workgroups of
// neighbouring rows share an L2; within an XCD the units of plane segment 0
// come first, then segment 1, ...  Segment s covers planes
// [s nz / nseg, (s + 1) nz / nseg).  With two rows per workgroup each lane
// computes one halo row's phi(t+1) per plane instead of two.
template <int RPW>
__global__ __launch_bounds__(RPW * kMarchCells) __attribute__((amdgpu_waves_per_eu(4))) void k1_march(
    const Scene sc, int q, int nseg, int cols8) {
  using C = Codec<ST_F32>;
  using R = double;
  __shared__ MarchLds<RPW> L;
  Geom gm = sc.g;
  const Geom &g = gm;
  const Derived &d = sc.d;
  const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
  const int seg = j / cols8, c = j - seg * cols8;
  const int U = (g.ny + RPW - 1) / RPW;
  const int ulo = (int)(((int64_t)xcd * U) >> 3), uhi = (int)(((int64_t)(xcd + 1) * U) >> 3);
  if (ulo + c >= uhi || seg >= nseg) return;  // the whole workgroup
  const int yb = (ulo + c) * RPW;             // the workgroup's first row
  const int z0 = (int)((int64_t)seg * g.nz / nseg), z1 = (int)((int64_t)(seg + 1) * g.nz / nseg);
  const int r = RPW == 1 ? 0 : (int)(threadIdx.x >> 8);  // own row of the lane
  const int x = (int)(threadIdx.x & (kMarchCells - 1));
  const int y = yb + r;
  const bool on = x < g.nx && y < g.ny;
  int box[6];
#pragma unroll
  for (int k = 0; k < 6; k++) box[k] = sc.ctrl->box_fused[k];
  q &= 1;
  const float *__restrict__ f2 = pop<ST_F32>(sc.f, q);
  const float *__restrict__ h2 = pop<ST_F32>(sc.h, q);
  float *__restrict__ fo = pop<ST_F32>(sc.f, q ^ 1);
  float *__restrict__ ho = pop<ST_F32>(sc.h, q ^ 1);
  const float *__restrict__ phi = par(sc.phi, q);
  float *__restrict__ phi_next = par(sc.phi, q ^ 1);
  const uint8_t *__restrict__ tag = sc.tag;
  // Rows 0 and ny - 1 are walls: nothing streams there and their neighbour
  // rows' phi(t+1) is not needed.
  const bool inner_row = on && y >= 1 && y <= g.ny - 2;

  // phi(t+1) of the lane's own row at plane z into slot z & 3 (and phi_next
  // when the plane is the workgroup's own); with keep_h its pulled h into
  // slot s3.
  auto own_phase = [&](int z, int s3, bool store, bool keep_h) {
    if (!on) return;
    float hv[Q7];
    bool fl;
    const float p = march_phase(g, box, tag, h2, phi, x, y, z, hv, fl);
    L.p[r][z & 3][x] = p;
    if (store) stb(phi_next, (uint32_t)(((int64_t)z * g.ny + y) * g.nx + x) * 4u, p);
    if (keep_h && fl) {
#pragma unroll
      for (int k = 0; k < Q7; k++) L.h[r][s3][k][x] = hv[k];
    }
  };
  // phi(t+1) of the halo rows at plane z (only where an inner own row needs it).
  auto halo_phase = [&](int z, int s3) {
    if (!inner_row) return;
    float hv[Q7];
    bool fl;
    if (RPW == 1) {
      L.hal[0][s3][x] = march_phase(g, box, tag, h2, phi, x, y - 1, z, hv, fl);
      L.hal[1][s3][x] = march_phase(g, box, tag, h2, phi, x, y + 1, z, hv, fl);
    } else {
      const int yy = r == 0 ? y - 1 : y + 1;
      L.hal[r][s3][x] = march_phase(g, box, tag, h2, phi, x, yy, z, hv, fl);
    }
  };
  // stream(t) + macro(t) + collide(t+1) of the lane's cell at plane z when it
  // is fluid and outside the band box (stream_collide_body's bulk path), with
  // the phi(t+1) stencil from LDS.
  auto stream_cell = [&](int z, int s3) {
    if (!inner_row || z < 1 || z > g.nz - 2 || x < 1 || x > g.nx - 2) return;
    if (in_box(box, x, y, z)) return;
    const int64_t i = ((int64_t)z * g.ny + y) * g.nx + x;
    const uint32_t bo4 = (uint32_t)i * 4u;
    const Phi7 v = load_phi7(g, phi, bo4);
    const R phiOld = v.c;
    const R rhoPhaseOld = R(d.rhoG) + phiOld * R(d.dRho);
    float gxo, gyo, gzo, lapo;
    phi_grad7<R>(v, phiOld, gxo, gyo, gzo, lapo);
    const uint32_t nonfluid = wall_mask(g, x, y, z);
    const RowAddr ra = row_addr(g, i, z * g.ny + y);
    double mom[6] = {0, 0, 0, 0, 0, 0};
    float fp[Q27];
    R pSumF, mx, my, mz;
    pull27<ST_F32>(g, i, ra, f2, nonfluid, 0u, sc.usx, sc.usy, sc.usz, rhoPhaseOld, false, x, y,
                   z, 0.0, 0.0, 0.0, mom, fp, pSumF, mx, my, mz);
    float hp[Q7];
#pragma unroll
    for (int k = 0; k < Q7; k++) hp[k] = L.h[r][s3][k][x];
    const int a = z & 3;
    Phi7 p1;
    p1.c = L.p[r][a][x];
    p1.px = L.p[r][a][x + 1];
    p1.mx = L.p[r][a][x - 1];
    if (RPW == 1) {
      p1.my = L.hal[0][s3][x];
      p1.py = L.hal[1][s3][x];
    } else {
      p1.my = r == 0 ? L.hal[0][s3][x] : L.p[0][a][x];
      p1.py = r == 0 ? L.p[RPW - 1][a][x] : L.hal[1][s3][x];
    }
    p1.pz = L.p[r][(z + 1) & 3][x];
    p1.mz = L.p[r][(z - 1) & 3][x];
    float *fob = fo + ra.base, *hob = ho + ra.base;
    finish_cell<true, ST_F32>(
        sc, i, false, phi_next, gxo, gyo, gzo, lapo, fp, hp, pSumF, mx, my, mz, 0, phi_next,
        [&](int k, float val, float base) {
          C::sto(fob, ra.relb + (uint32_t)k * (uint32_t)g.pb * 4u, val, base);
        },
        [&](int k, float val, float base) {
          C::sto(hob, ra.relb + (uint32_t)k * (uint32_t)g.pb * 4u, val, base);
        },
        &p1);
  };

  int s3 = z0 % 3;  // slot of plane z in hal and h
  if (z0 >= 1) own_phase(z0 - 1, 0, false, false);
  own_phase(z0, s3, true, true);
  halo_phase(z0, s3);
  for (int z = z0; z < z1; z++) {
    const int n3 = s3 == 2 ? 0 : s3 + 1;
    // The ~100 uniform pull offsets derived from the geometry are cheaper to
    // recompute per plane than to keep live through the loop (scalar
    // registers would spill): the geometry is made opaque per iteration.
    asm volatile("" : "+s"(gm.nx), "+s"(gm.ny), "+s"(gm.nxny), "+s"(gm.pa), "+s"(gm.pb));
    if (z + 1 < g.nz) {
      // Plane z1 belongs to the next segment: only the own rows' phi(t+1) is
      // needed here (the stencil of plane z1 - 1).
      const bool mine = z + 1 < z1;
      own_phase(z + 1, n3, mine, mine);
      if (mine) halo_phase(z + 1, n3);
    }
    lds_barrier();
    stream_cell(z, s3);
    s3 = n3;
  }
}

// Whether a scene can march: rows of at most kMarchCells cells.
inline bool march_fits(const Geom &g) { return g.nx >= 3 && g.nx <= kMarchCells && g.ny >= 3; }

// Launch shape of the march with RPW rows per workgroup: row units per XCD
// (cols8) and plane segments (nseg).  Unless nseg is given, segments of about
// eight planes, and at least 4096 / RPW waves (sixteen per CU on 256 CUs):
// short segments cost a little redundant phase work at their ends (two extra
// planes of phi(t+1) per segment) but keep many independent workgroups in
// flight (256^3 on MI355X, two rows per workgroup: 1.131 ms per step with 8
// segments, 1.093-1.102 with 32 or 48, 1.101-1.110 with 64-128).
inline void march_shape(const Geom &g, int rpw, int nseg_req, int &cols8, int &nseg) {
  cols8 = ((g.ny + rpw - 1) / rpw + 7) / 8;
  const int wgs = 1024 / rpw;
  const int by_waves = (wgs + 8 * cols8 - 1) / (8 * cols8);
  const int by_planes = (g.nz + 7) / 8;
  nseg = nseg_req > 0 ? nseg_req : (by_planes > by_waves ? by_planes : by_waves);
  if (nseg < 1) nseg = 1;
  if (nseg > g.nz) nseg = g.nz;
}

}  // namespace lbm
