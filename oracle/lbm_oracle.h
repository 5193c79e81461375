/*
 * CPU oracle for the coupled LBM step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference simulator's solver (src/sim.js), rigid
 * sphere (src/solid.js) and lattice tables (src/lattices.js).  It exists to check
 * the HIP library: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker or as the timed CPU
 * baseline.  Nothing in the product links it.
 *
 * Arithmetic follows the reference exactly: every expression is evaluated in
 * IEEE double in the reference's left-to-right order, and every array store
 * rounds to the storage type (float32, like the reference's Float32Array).
 * Build with -ffp-contract=off so no multiply-add is fused.
 *
 * Storage type: ostore_t is float by default.  Compiled with -DORACLE_FP64 the
 * populations and macroscopic fields are stored in double instead, giving the
 * pure-FP64 oracle used to measure the error of reduced-precision storage.
 *
 * Layout: the reference's own (array-of-structures f[i*27+k], h[i*7+k], x
 * fastest, i = (z*ny + y)*nx + x).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

#ifdef ORACLE_FP64
typedef double ostore_t;
#else
typedef float ostore_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Field set of class Sim (sim.js:30-96). */
typedef struct osim {
  int nx, ny, nz, n;
  ostore_t *f, *f2, *h, *h2;
  ostore_t *rho, *ux, *uy, *uz, *phi, *phiPrev;
  uint8_t *tag, *tagPrev;
  ostore_t *usx, *usy, *usz;
  ostore_t *gx, *gy, *gz, *lap;
  ostore_t *bodyFx, *bodyFy, *bodyFz;
  int32_t neighOffset27[27], neighOffset7[7];
  double tau, tauPhi, sigma, gravity, rhoL, rhoG, W, mobility;
} osim;

/* Field set of class RigidSphere (solid.js:9-37), all doubles; dragging is 0/1. */
typedef struct osphere {
  double nx, ny, nz;
  double cx, cy, cz, r;
  double vx, vy, vz;
  double wx, wy, wz;
  double density, volume, mass, I;
  double dragging, dragTargetX, dragTargetY, dragTargetZ;
  double Fx, Fy, Fz, Tx, Ty, Tz;
} osphere;

/* lattices.js */
int oracle_tables(int8_t *ex27, int8_t *ey27, int8_t *ez27, float *w27, int8_t *opp27,
                  int8_t *ex7, int8_t *ey7, int8_t *ez7, float *w7, int8_t *opp7);

/* sim.js */
osim *oracle_sim_create(int nx, int ny, int nz);
void oracle_sim_destroy(osim *s);
int oracle_storage_bytes(void);
void oracle_init_flat(osim *s, double waterlineY);
void oracle_init_waterline_map(osim *s, const double *yw_xz);
void oracle_seed_hydrostatic(osim *s);
void oracle_compute_phi_gradients(osim *s);
void oracle_retag_cells(osim *s, const osphere *solid);
void oracle_handle_fresh_cells(osim *s);
void oracle_collide_phase(osim *s);
void oracle_stream_phase(osim *s);
void oracle_collide_hydro(osim *s);
void oracle_stream_hydro_and_macro(osim *s, osphere *solid);
void oracle_sim_step(osim *s, osphere *solid);
void oracle_pack_volume(const osim *s, uint8_t *out);

/* solid.js */
void oracle_sphere_init(osphere *b, int nx, int ny, int nz, double cx, double cy, double cz,
                        double r, double density);
void oracle_sphere_recompute_mass(osphere *b);
int oracle_sphere_contains(const osphere *b, int x, int y, int z);
void oracle_sphere_surface_velocity(const osphere *b, int x, int y, int z, double *out);
void oracle_sphere_apply_impulse_at_cell(osphere *b, int x, int y, int z,
                                         double px, double py, double pz);
void oracle_sphere_absorb_dead_cells(osphere *b, const osim *s);
void oracle_sphere_step(osphere *b, double dt, double gravity);

/* Driver contract (worker.js:96-100): sim.step; absorbDeadCells; body step. */
void oracle_coupled_steps(osim *s, osphere *b, int nsteps);

#ifdef __cplusplus
}
#endif
#endif
