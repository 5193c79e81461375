// Device-side types shared by every kernel of liblbm.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

namespace lbm {

// Grid geometry.  Cell arrays (phi, rho, tags, ...) hold cell i = (z*ny + y)*nx + x
// at i.  S = n rounded up to 64 cells (256 bytes) is the plane stride of the
// 16-bit population codecs, which keep population k of cell i at k*S + i.
//
// Float32 populations: one buffer per step parity holds the 27 f directions
// and then the 7 h directions (the h pointer is the f pointer plus Q27*pb);
// population k of cell i (row r = i / nx = z*ny + y) is element
//   i + r*pa + k*pb
// of its pointer.  Two layouts:
//  * row blocks (pa = 33 nx, pb = nx): every grid row's 34 directions side by
//    side, element (r*34 + k)*nx + x.  A lane's pulls land in the 34 row chunks
//    of three neighbouring row blocks instead of 34 planes a whole grid apart:
//    the 34-direction pull-and-store pattern runs at 5.6 TB/s against 5.25 for
//    planes at 256^3 on MI355X (5.6 against 5.1 at 384^3);
//  * planes (pa = 0, pb = S): direction k of every cell, then the next.
//    Narrower grids run faster so (64 scenes of 32..96 cells per axis stepped
//    together: 10.5 k MLUPS in planes against 8.96 k in row blocks; served
//    scenes of 32..192 cells per axis: 11.4 k against 10.5 k MLUPS with row
//    blocks from nx >= 256 against nx >= 128).
// lbm_create chooses row blocks when nx >= kRowLayoutMinNx (or when planes
// would not fit the fused kernels' 32-bit offsets, f32_fused_fits).  Either way
// a direction's source element is a uniform offset from the cell's own.
struct Geom {
  int nx, ny, nz, nxny;
  int64_t n, S;
  int64_t pa, pb;  // float32 population layout (above)
};

constexpr int kRowQ = 27 + 7;
constexpr int kRowLayoutMinNx = 256;
enum : int { LAYOUT_PLANES = 0, LAYOUT_ROWS = 1 };
inline void set_layout(Geom &g, int layout) {
  g.pa = layout == LAYOUT_ROWS ? (int64_t)(kRowQ - 1) * g.nx : 0;
  g.pb = layout == LAYOUT_ROWS ? (int64_t)g.nx : g.S;
}
// Element of direction 0 of cell i in a float32 population buffer.
__host__ __device__ inline int64_t pop_base(const Geom &g, int64_t i) {
  return g.pa ? i + (i / g.nx) * g.pa : i;
}
// Offset from a cell's element of direction k to its source's element of the
// same direction: the source is x - ex, y - ey, z - ez, that is (ez*ny + ey)
// rows and ex cells before.
__host__ __device__ inline int64_t pop_src_delta(const Geom &g, int ex, int ey, int ez) {
  return ((int64_t)ez * g.ny + ey) * (g.nx + g.pa) + ex;
}

// Per-step constants, computed on the host from lbm_params with the same
// double-precision expressions the reference evaluates at the top of each pass.
struct Derived {
  // collideHydro / streamHydroAndMacro (src/sim.js:410-416, 490-492)
  double invTau, beta, kappa, dRho, A, rhoG, gravity;
  // collidePhase (src/sim.js:327-332)
  double invTauPhi, sharpC, sharpFac;
};

// The same constants rounded to float, for the FP32 arithmetic of the 16-bit
// codecs (half the scalar registers of the doubles, no conversions per use).
struct DerivedF {
  float invTau, beta, kappa, dRho, A, rhoG, gravity;
  float invTauPhi, sharpC, sharpFac;
};
inline DerivedF derive_f(const Derived &d) {
  return DerivedF{(float)d.invTau, (float)d.beta,      (float)d.kappa,  (float)d.dRho,
                  (float)d.A,      (float)d.rhoG,      (float)d.gravity, (float)d.invTauPhi,
                  (float)d.sharpC, (float)d.sharpFac};
}

// Device control block: list counters, error flags, the device copy of the
// sphere and the band boxes of the fused step.
struct Ctrl {
  int n_fresh, n_dead, n_solid, n_links;
  int err;         // nonzero: a device-side check failed (see kErr*)
  int n_partials;  // blocks of the fused stream kernel that touched the body
  int box_band[6];   // x0,x1,y0,y1,z0,z1 (inclusive) of the band handled by the band kernels
  int box_fused[6];  // band box that the next fused launch must skip
  lbm_sphere sphere; // device sphere for the fused coupled step
  double F_links[6]; // link momentum of the last fused step (Fx,Fy,Fz,Tx,Ty,Tz), diagnostics
};

constexpr int kErrListOverflow = 1;
constexpr int kErrBandEscape = 2;
constexpr int kErrFreshCapacity = 4;
constexpr int kErrBandTiles = 8;  // a band-tile launch did not cover every band-row tile
constexpr int kErrLinksUnsorted = 16;  // link list too long for k_sort_apply_links

// Tiles of 256 consecutive cells in the x-fastest order (four waves of 64),
// launched on a 1-D grid in XCD-aware order.
//  * Shape: a tile covers one row of a 256-wide grid, several rows of a narrow
//    one, and leaves no lane idle except in the last tile.  Each population
//    plane is read and written in 1 KB runs per workgroup; with 34 planes in
//    flight, long runs keep HBM pages open far better than 256 B runs
//    (measured on MI355X: 4.4 vs 3.9 TB/s for the 34-plane pull pattern).
//  * Order: the hardware hands consecutive workgroups to the eight XCDs in
//    turn, and each XCD has its own L2; mapping physical workgroup p to tile
//    (p % 8) * (T8 / 8) + p / 8 gives every XCD a contiguous run of tiles (a
//    slab of z-planes), so tiles that share halo rows run on the same L2.
struct TileMap {
  int nx, nxny;  // row length and plane size in cells
  int T, T8;     // tiles, and rounded up to a multiple of 8 (the grid)
};

constexpr int kTileShift = 8;  // 256 cells per tile: one cell per lane
// D of the fused float32 pulls: the largest distance of a source element below
// a cell's own element of the same direction ((ny + 1) rows and one cell).
__host__ __device__ inline int64_t pull_span(const Geom &g) {
  return (int64_t)(g.ny + 1) * (g.nx + g.pa) + 1;
}
// The fused float32 kernels address every population of a wave by a 32-bit
// byte offset from one base per buffer: a lane's offset (within three rows),
// up to 33 direction strides and twice D.  True when that stays below 2^31
// bytes (so element offsets also fit 32 bits).
inline bool f32_fused_fits(const Geom &g) {
  const int64_t span = 64 + 3 * (g.nx + g.pa) + (kRowQ - 1) * g.pb + 2 * pull_span(g);
  return span * 4 < ((int64_t)1 << 31);
}
// The layout lbm_create gives a grid (g.S set): row blocks from nx >= 256, or
// when planes would not fit the fused kernels' offsets.
inline void choose_layout(Geom &g) {
  set_layout(g, LAYOUT_PLANES);
  if (g.nx >= kRowLayoutMinNx || !f32_fused_fits(g)) set_layout(g, LAYOUT_ROWS);
}

inline TileMap make_tile_map(int nx, int ny, int nz) {
  TileMap m;
  m.nx = nx;
  m.nxny = nx * ny;
  const int64_t n = (int64_t)m.nxny * nz;
  m.T = (int)((n + (1 << kTileShift) - 1) >> kTileShift);
  m.T8 = (m.T + 7) / 8 * 8;
  return m;
}

// Tile of launch-local workgroup lb of a grid of T8 workgroups (XCD-aware
// order); false for the padding workgroups past T.
__device__ __forceinline__ bool map_tile(int T, int T8, int lb, int &tile) {
  tile = (lb & 7) * (T8 >> 3) + (lb >> 3);
  return tile < T;
}

// Quotient and remainder of a / b for 0 <= a < 2^31 and 1 <= b, when the
// quotient is below 2^20: the float estimate is then within one of it, and
// one correction each way makes it exact.
__device__ __forceinline__ int small_divmod(int a, int b, int &rem) {
  int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
  int r = a - q * b;
  if (r < 0) {
    q -= 1;
    r += b;
  }
  if (r >= b) {
    q += 1;
    r -= b;
  }
  rem = r;
  return q;
}

// Cell coordinates of index i (grids of fewer than 2^31 cells, axes below 2^20).
__device__ __forceinline__ void cell_xyz(int nx, int nxny, int64_t i, int &x, int &y, int &z) {
  int r;
  z = small_divmod((int)i, nxny, r);
  y = small_divmod(r, nx, x);
}

// True when tile t holds a cell of the band box's rows: in some box plane z the
// tile lies between the tiles of that plane's first and last box cell.  Every
// cell of the box lies in such a tile.
__host__ __device__ inline bool tile_in_box_rows(const TileMap &m, const int box[6], int t) {
  const int64_t c0 = (int64_t)t << kTileShift, c1 = c0 + (1 << kTileShift) - 1;
  int za = (int)(c0 / m.nxny), zb = (int)(c1 / m.nxny);
  za = za < box[4] ? box[4] : za;
  zb = zb > box[5] ? box[5] : zb;
  for (int z = za; z <= zb; z++) {
    const int64_t p = (int64_t)z * m.nxny;
    const int lo = (int)((p + (int64_t)box[2] * m.nx + box[0]) >> kTileShift);
    const int hi = (int)((p + (int64_t)box[3] * m.nx + box[1]) >> kTileShift);
    if (t >= lo && t <= hi) return true;
  }
  return false;
}

__device__ __forceinline__ int64_t cell_index(const Geom &g, int x, int y, int z) {
  return ((int64_t)z * g.ny + y) * g.nx + x;
}

__device__ __forceinline__ bool interior(const Geom &g, int x, int y, int z) {
  return x >= 1 && x < g.nx - 1 && y >= 1 && y < g.ny - 1 && z >= 1 && z < g.nz - 1;
}

// Pull offset of direction k: the source of cell i is i + off = i - e_k
// (src/sim.js:71-78).
template <int K>
__device__ __forceinline__ int64_t pull_off27(const Geom &g) {
  return -((int64_t)EZ27[K] * g.nxny + (int64_t)EY27[K] * g.nx + EX27[K]);
}

}  // namespace lbm
