/*
 * liblbm -- C ABI of the MI355X (gfx950) coupled lattice-Boltzmann step.
 *
 * One coupled step is the reference simulator's driver sequence
 * (src/worker.js:96-100, test/bench.mjs:22-26):
 *     sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity)
 * where Sim.step (src/sim.js:202-210) runs retag -> fresh cells -> phi
 * gradients -> phase collide -> phase stream -> hydro collide -> hydro stream
 * with macroscopic update.
 *
 * Every declaration below names the reference interface it replaces.  Plain C
 * types only; all state lives in device memory owned by an lbm_sim handle.
 * Functions return LBM_OK (0) or a negative LBM_ERR_* code, with a message
 * available from lbm_last_error().  Calls on one handle must not overlap; each
 * handle owns its own HIP stream.  Calls are synchronous unless stated.
 *
 * Field layout at the boundary is the reference's: cell i = (z*ny + y)*nx + x
 * (x fastest), f as f[i*27 + k] and h as h[i*7 + k] in the direction order of
 * src/lattices.js:34-49 and :79-84, float32 values, tags as uint8 (0 fluid,
 * 1 solid, 2 wall).  The device keeps populations as structure-of-arrays.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_ARG (-1)      /* bad argument (null handle, bad size, bad field id) */
#define LBM_ERR_OOM (-2)      /* device or host allocation failed */
#define LBM_ERR_HIP (-3)      /* a HIP runtime call failed */
#define LBM_ERR_STATE (-4)    /* call not valid in the current state */
#define LBM_ERR_CAPACITY (-5) /* a bounded list overflowed */
#define LBM_ERR_NODEVICE (-6) /* no usable GPU */

/* Cell tags, src/lattices.js:102-104. */
#define LBM_TAG_FLUID 0
#define LBM_TAG_SOLID 1
#define LBM_TAG_WALL 2

typedef struct lbm_sim lbm_sim;

/* Physics parameters: the public scalar fields of class Sim (src/sim.js:88-95).
 * They are read at every pass, so they may change between steps
 * (src/worker.js:139-145).  mobility is carried but unused, as in the
 * reference: the phase mobility is cs7^2 (tauPhi - 1/2) (src/sim.js:329). */
typedef struct lbm_params {
  double tau, tauPhi, sigma, gravity, rhoL, rhoG, W, mobility;
} lbm_params;

/* The fields of class RigidSphere (src/solid.js:9-37), all FP64, in this order.
 * dragging is 0 or 1.  Fx..Tz are the momentum-exchange accumulators. */
typedef struct lbm_sphere {
  double nx, ny, nz;
  double cx, cy, cz, r;
  double vx, vy, vz;
  double wx, wy, wz;
  double density, volume, mass, I;
  double dragging, dragTargetX, dragTargetY, dragTargetZ;
  double Fx, Fy, Fz, Tx, Ty, Tz;
} lbm_sphere;

/* Field identifiers for lbm_download / lbm_upload: the typed arrays of class
 * Sim (src/sim.js:35-64). */
enum lbm_field {
  LBM_FIELD_F = 0,  /* n*27 float */
  LBM_FIELD_F2,     /* n*27 float */
  LBM_FIELD_H,      /* n*7 float */
  LBM_FIELD_H2,     /* n*7 float */
  LBM_FIELD_RHO,    /* n float: the pressure proxy sum(f) */
  LBM_FIELD_UX,
  LBM_FIELD_UY,
  LBM_FIELD_UZ,
  LBM_FIELD_PHI,
  LBM_FIELD_PHIPREV,
  LBM_FIELD_TAG,    /* n uint8 */
  LBM_FIELD_TAGPREV,/* n uint8 */
  LBM_FIELD_USX,
  LBM_FIELD_USY,
  LBM_FIELD_USZ,
  LBM_FIELD_GX,
  LBM_FIELD_GY,
  LBM_FIELD_GZ,
  LBM_FIELD_LAP,
  LBM_FIELD_BODYFX,
  LBM_FIELD_BODYFY,
  LBM_FIELD_BODYFZ,
  LBM_FIELD_COUNT
};

/* ------------------------------------------------------------------ lifetime */

/* new Sim(nx, ny, nz) (src/sim.js:30-96): allocates the device state and sets
 * the reference's default parameters.  The state is zero until an init call.
 * At most 2^30 cells; lbm_step with float32 storage addresses two planes of
 * populations with 32-bit offsets, which takes grids whose planes (nx * ny)
 * hold up to about 7.9 M cells (any nz; the reference's typed arrays end near
 * 2^26 cells in all) and refuses larger planes with LBM_ERR_ARG. */
int lbm_create(int nx, int ny, int nz, lbm_sim **out);
void lbm_destroy(lbm_sim *sim);
/* Message for the last failed call on this handle (or the last failed
 * lbm_create when sim is NULL). */
const char *lbm_last_error(const lbm_sim *sim);
/* Grid size; n = nx*ny*nz. */
int lbm_dims(const lbm_sim *sim, int *nx, int *ny, int *nz);

/* Assignments to sim.tau, sim.tauPhi, ... (src/sim.js:88-95). */
int lbm_set_params(lbm_sim *sim, const lbm_params *p);
int lbm_get_params(const lbm_sim *sim, lbm_params *p);

/* ------------------------------------------------------------------ init */

/* Sim.initFlat(waterlineY) (src/sim.js:112-134), including its call of
 * seedHydrostatic. */
int lbm_init_flat(lbm_sim *sim, double waterline_y);
/* initFlat with one waterline per (x,z) column, yw[z*nx + x]; a constant map
 * gives exactly lbm_init_flat.  No reference counterpart: it seeds perturbed
 * interfaces. */
int lbm_init_waterline_map(lbm_sim *sim, const double *yw);
/* Sim.seedHydrostatic() (src/sim.js:144-173). */
int lbm_seed_hydrostatic(lbm_sim *sim);

/* ------------------------------------------------------------------ passes */
/* The seven passes of Sim.step, one entry each, with the reference's array
 * semantics (f -> f2 -> f, h -> h2 -> h, stored gradients).  `solid` may be
 * NULL, as sim.step(null) allows (src/sim.js:233, 527). */

/* Sim.retagCells(solid) (src/sim.js:214-244) for a RigidSphere. */
int lbm_retag_cells(lbm_sim *sim, const lbm_sphere *solid);
/* Sim.retagCells(solid) for any other solid: the caller evaluated
 * solid.contains (mask[i] != 0 on interior cells) and solid.surfaceVelocity
 * (us[3*i+c], stored as float like the reference) on the host. */
int lbm_retag_cells_mask(lbm_sim *sim, const uint8_t *mask, const float *us);
/* Sim.handleFreshCells(solid) (src/sim.js:257-310); uses the fresh cells
 * found by the last retag. */
int lbm_handle_fresh_cells(lbm_sim *sim);
/* Sim.computePhiGradients() (src/sim.js:176-199). */
int lbm_compute_phi_gradients(lbm_sim *sim);
/* Sim.collidePhase() (src/sim.js:324-364). */
int lbm_collide_phase(lbm_sim *sim);
/* Sim.streamPhase() (src/sim.js:376-397). */
int lbm_stream_phase(lbm_sim *sim);
/* Sim.collideHydro() (src/sim.js:406-464). */
int lbm_collide_hydro(lbm_sim *sim);
/* Sim.streamHydroAndMacro(solid) (src/sim.js:484-582).  With a sphere, every
 * fluid->solid link's impulse is added to solid->F/T through
 * lbm_sphere_apply_impulse_at_cell in the reference's cell-then-direction
 * order, so the accumulators match the reference bit for bit. */
int lbm_stream_hydro_macro(lbm_sim *sim, lbm_sphere *solid);

/* One momentum-exchange record of lbm_stream_hydro_macro_links. */
typedef struct lbm_link {
  int32_t x, y, z;    /* solid (source) cell, the applyImpulseAtCell position */
  int32_t k;          /* lattice direction of the fluid population */
  double px, py, pz;  /* impulse handed to the body */
} lbm_link;
/* streamHydroAndMacro for a solid that is not a RigidSphere: instead of
 * accumulating into a sphere, the handle keeps every link's applyImpulseAtCell
 * call in reference order (src/sim.js:527-536) so the host can replay them on
 * its own object.  *count receives the number of links; lbm_get_links copies
 * them out. */
int lbm_stream_hydro_macro_links(lbm_sim *sim, int *count);
/* Copies min(count, cap) links kept by the last lbm_stream_hydro_macro_links. */
int lbm_get_links(lbm_sim *sim, lbm_link *out, int cap);

/* Sim.step(solid) (src/sim.js:202-210): the seven passes above.  With a
 * sphere, the step's link impulses are applied on the device, sorted by
 * (cell, direction) and added one after the other, and *solid is read back
 * once: the same bits as the host replay of lbm_stream_hydro_macro. */
int lbm_sim_step(lbm_sim *sim, lbm_sphere *solid);

/* RigidSphere.absorbDeadCells(sim) (src/solid.js:87-101): adds each dead cell's
 * impulse (cells that became solid at the last retag), in index order, on the
 * device; *solid is uploaded and read back. */
int lbm_absorb_dead_cells(lbm_sim *sim, lbm_sphere *solid);

/* ------------------------------------------------------------------ coupled */

/* nsteps coupled steps (src/worker.js:96-100): per step sim.step(solid);
 * solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity).  This is the fast
 * path: the sphere is integrated on the device, passes are fused, and the host
 * synchronises once per call.  Field results equal lbm_sim_step's; the force
 * sums use a fixed parallel order, so the sphere state can differ from the
 * per-pass path in the last bits.  solid may be NULL (no body). */
int lbm_step(lbm_sim *sim, lbm_sphere *solid, int nsteps);

/* nsteps coupled steps like lbm_step, ending live: the call stops after the
 * last step's sphere update and defers that step's band work (the next
 * step's retag, fresh cells and collide near the sphere) to the next call.  No
 * reference counterpart.  A call that starts from the reference's state runs
 * a prologue (a retag, fresh-cell pass and collide of every cell) and one that
 * ends in it a stream-only last step; live calls skip both, so consecutive
 * lbm_step_live calls cost what one long call costs -- a frame loop that steps
 * two substeps per frame (src/worker.js:94-103) pays no per-call passes.
 * Between live calls phi, phiPrev, the tags, the surface velocities, the
 * gradients, the fresh/dead lists and the sphere are the reference's after
 * those steps (lbm_pack_volume works); the populations, rho and u are not, and
 * every entry point that needs them (their download, any upload, the per-pass
 * calls, lbm_set_storage, lbm_mass, lbm_check_fields) returns LBM_ERR_STATE
 * until an lbm_step call of at least one step has ended in the reference's
 * state.  lbm_init_flat / lbm_init_waterline_map start over. */
int lbm_step_live(lbm_sim *sim, lbm_sphere *solid, int nsteps);

/* lbm_step that allocates no device memory.  No reference counterpart: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and results as
 * lbm_step, bit for bit; the fused step's link-momentum partials go to a
 * library-wide device array of 16 MiB held on each GPU until the process ends
 * (grids of at most 89,477,120 cells, 384^3 included; LBM_ERR_CAPACITY beyond),
 * and steps are launched one by one (no launch graph).  The handle's
 * own buffers must already exist (they do after lbm_create, and after
 * lbm_set_storage for 16-bit storage); timing must be off.  Calls that overlap
 * on one GPU share that array and race: run them one at a time per GPU. */
int lbm_step_noalloc(lbm_sim *sim, lbm_sphere *solid, int nsteps);

/* ------------------------------------------------------------------ storage */
/* Population storage of the fused step.  No reference counterpart (the
 * reference stores float32, src/sim.js:35-38).  LBM_STORAGE_F32 is the
 * reference-exact form.  LBM_STORAGE_F16 / LBM_STORAGE_BF16 store every
 * population as a 16-bit deviation from its direction's baseline (f_k - w_k,
 * h_k - w7_k phi) and compute the collide in FP32; the momentum-exchange
 * force/torque sums stay FP64.  Results then differ from the reference by the
 * storage rounding; the other fields (phi, rho, u, tags) stay float32.  The
 * float32 populations are decoded on demand (field reads and writes, per-pass
 * calls); after a compressed lbm_step only f and h are defined, not f2/h2. */
#define LBM_STORAGE_F32 0
#define LBM_STORAGE_F16 1
#define LBM_STORAGE_BF16 2
int lbm_set_storage(lbm_sim *sim, int kind);
int lbm_get_storage(const lbm_sim *sim);
/* Global mass check: FP64 sums of phi and of rho (sum f) over fluid cells,
 * added in a fixed order.  No reference counterpart. */
int lbm_mass(lbm_sim *sim, double *phi_sum, double *rho_sum);
/* Opt-in state check, off the hot path: *all_finite = 1 when phi, rho and u are
 * finite at every fluid cell, *max_u the largest |u| there.  The reference has
 * no error paths; its callers scan for NaN (test/smoke.mjs:33-40). */
int lbm_check_fields(lbm_sim *sim, int *all_finite, double *max_u);

/* ------------------------------------------------------------------ batches */
/* Many simulations stepped together.  No reference counterpart: the reference
 * runs one Sim per Web Worker (src/worker.js:11-13); a batch steps any number
 * of independent Sims (each with its own size, parameters and sphere) with
 * launches that cover all of them, so small scenes fill the GPU together.
 * Every member's results are those of lbm_step on it alone, bit for bit.
 * Members of different storage kinds step in separate launches. */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(lbm_batch **out);
void lbm_batch_destroy(lbm_batch *batch);
/* Message for the last failed call on this batch. */
const char *lbm_batch_last_error(const lbm_batch *batch);
/* Adds sim (solid may be NULL); the batch keeps both pointers and updates
 * *solid after every lbm_batch_step.  Returns the member index (>= 0). */
int lbm_batch_add(lbm_batch *batch, lbm_sim *sim, lbm_sphere *solid);
int lbm_batch_remove(lbm_batch *batch, lbm_sim *sim);
int lbm_batch_size(const lbm_batch *batch);
/* nsteps coupled steps of every member (lbm_step semantics). */
int lbm_batch_step(lbm_batch *batch, int nsteps);
/* Per-kernel HIP event timing of lbm_batch_step, as lbm_set_timing. */
int lbm_batch_set_timing(lbm_batch *batch, int on);
int lbm_batch_kernel_time(lbm_batch *batch, const char *name, double *total_ms, int64_t *launches);
/* Launch graphs of lbm_batch_step, as lbm_set_graph_steps / lbm_graph_counts.
 * A set's graph reads the scene table from device memory, so it is replayed
 * for any members whose launch grids match the captured ones. */
int lbm_batch_set_graph_steps(lbm_batch *batch, int steps);
int lbm_batch_graph_counts(const lbm_batch *batch, int64_t *captures, int64_t *replays);
/* Bulk/band-row split of the batch's stream kernel, as lbm_set_stream_split:
 * 1 split, 0 one launch, -1 (default) by storage kind (split for the 16-bit
 * codecs, one launch for float32). */
int lbm_batch_set_stream_split(lbm_batch *batch, int on);
/* Cells per lane of the batch's 16-bit stream kernel, as lbm_set_stream_lanes. */
int lbm_batch_set_stream_lanes(lbm_batch *batch, int cells_per_lane);

/* Launch-shape autotuner.  No reference counterpart.  Times every candidate
 * launch configuration of the batch -- cells per lane of the 16-bit stream
 * kernel (when a member stores 16-bit populations), the stream kernel's
 * bulk/band-row split (lbm_set_stream_split, lbm_batch_set_stream_split),
 * for one float32 simulation whose rows fit it the row march in three
 * workgroup shapes (lbm_set_march_shape: two rows and the default segments,
 * one row, two rows and half the segments), and steps per captured graph
 * (0, 2, 8, 32, those below trial_steps - 1) -- over trial_steps real
 * coupled steps each, after an untimed lead-in of 2 steps, keeps the fastest
 * and reports it.  Every candidate computes the same state, so the trials
 * simply advance the members: after the call each has advanced by
 * candidates * (trial_steps + 2) steps, exactly as lbm_batch_step would have
 * moved it. */
typedef struct {
  int cells_per_lane;  /* 1 or 2 */
  int graph_steps;     /* 0: kernels launched one by one */
  double ms_per_step;  /* of the chosen configuration */
  int candidates;      /* configurations timed */
  int stream_split;    /* stream launch shape of the choice (1 split, 0 one launch, 2 march) */
  int march_rows;      /* rows per workgroup of the row march (1 or 2) */
  int march_segments;  /* plane segments of the row march (0: chosen per grid) */
} lbm_tune_result;
int lbm_batch_autotune(lbm_batch *batch, int trial_steps, lbm_tune_result *out);
/* The same for one simulation's lbm_step. */
int lbm_autotune(lbm_sim *sim, lbm_sphere *solid, int trial_steps, lbm_tune_result *out);

/* ------------------------------------------------------------------ serving */
/* A request loop that keeps many scenes resident and steps them with
 * continuous batching.  No reference counterpart as an API: it serves the
 * frame protocol of src/worker.js:94-165 (an RG8 phi+mask volume, as
 * Sim.packVolume writes it, plus the sphere state, every few substeps) for
 * many independent scenes at once.  Requests wait in arrival order until their
 * state fits the device-memory budget and an active slot is free; every active
 * scene advances in the same launches; a frame is produced every frame_every
 * steps and at the end of the step budget. */
typedef struct {
  int nx, ny, nz;
  lbm_params params;
  const double *waterline; /* nz*nx interface heights for lbm_init_waterline_map, or NULL */
  double waterline_y;      /* lbm_init_flat height when waterline is NULL */
  lbm_sphere sphere;       /* initial sphere (lbm_sphere_init) when has_sphere */
  int has_sphere;
  int steps;               /* step budget */
  int frame_every;         /* steps between frames */
  int storage;             /* LBM_STORAGE_* */
} lbm_request;

typedef struct {
  int64_t request;   /* id from lbm_server_submit */
  int status;        /* LBM_OK, or the error that ended the request */
  int index;         /* frame number within the request */
  int step;          /* steps done when the frame was packed */
  int last;          /* 1 on the request's final frame */
  int nx, ny, nz;
  int64_t bytes;     /* 2 * nx * ny * nz */
  double t_submit;   /* server clock (seconds) at submission */
  double t_ready;    /* server clock when the frame was ready */
  double latency;    /* t_ready - previous frame's t_ready (t_submit for the first) */
  lbm_sphere sphere; /* sphere state at the frame */
} lbm_frame_info;

typedef struct {
  int64_t submitted, admitted, completed, rejected;
  int active, pending, max_active_seen, queued_frames;
  int64_t frames, ticks;
  int64_t lattice_updates; /* cell updates done (all cells of every stepped scene) */
  double busy_seconds;     /* time inside batched steps */
  double uptime_seconds;
  int64_t bytes_in_use, budget;
} lbm_server_stats;

typedef struct lbm_server lbm_server;
/* Device bytes lbm_create + lbm_set_storage take for a grid (admission). */
int64_t lbm_state_bytes(int nx, int ny, int nz, int storage);
/* hbm_budget: device bytes the resident scenes may take; max_active: active
 * scene cap; frame_queue_bytes: unread frame bytes before the loop waits for
 * the reader (0: 1 GiB).  Starts the loop thread.  The scenes' device memory
 * comes from a memory pool of the server's own, which keeps what finished
 * scenes free for the next ones until lbm_server_destroy. */
int lbm_server_create(int64_t hbm_budget, int max_active, int64_t frame_queue_bytes,
                      lbm_server **out);
/* Stops the loop and frees every resident scene. */
void lbm_server_destroy(lbm_server *server);
const char *lbm_server_last_error(lbm_server *server);
/* Queues a request (copied); *id receives its id.  LBM_ERR_CAPACITY when its
 * state alone exceeds the budget.  Thread-safe. */
int lbm_server_submit(lbm_server *server, const lbm_request *request, int64_t *id);
/* Pops the oldest frame, waiting up to timeout_ms: returns 1 with *info and the
 * 2n RG8 bytes in rg8 (cap bytes available; NULL to drop them), 0 on timeout,
 * negative on error.  LBM_ERR_CAPACITY when cap is too small: the frame stays
 * queued and info->bytes gives the size needed.  Thread-safe. */
int lbm_server_next_frame(lbm_server *server, int timeout_ms, lbm_frame_info *info, uint8_t *rg8,
                          int64_t cap);
int lbm_server_stats_get(lbm_server *server, lbm_server_stats *out);
/* Seconds since the server was created (the clock of the frame times). */
double lbm_server_clock(lbm_server *server);
/* paused != 0: no scene joins or steps after the tick in progress (requests
 * still queue and are prepared); 0: resume.  Lets a caller queue a burst at
 * once. */
int lbm_server_pause(lbm_server *server, int paused);

/* ------------------------------------------------------------------ fields */

/* Number of elements of a field (n, n*27 or n*7). */
int64_t lbm_field_count(const lbm_sim *sim, int field);
/* Reads of the Sim typed arrays (sim.phi[i] etc.), in the reference layout. */
int lbm_download(lbm_sim *sim, int field, void *host);
/* Writes to the Sim typed arrays (sim.phi[i] = ..., src/worker.js:61-67).  A
 * write to bodyF arms the external force for the next step; the step clears it
 * (src/sim.js:581). */
int lbm_upload(lbm_sim *sim, int field, const void *host);
/* Sim.packVolume(out) (src/sim.js:586-600): out has 2n bytes, R = phi*255
 * truncated, G = 255 on solid cells. */
int lbm_pack_volume(lbm_sim *sim, uint8_t *out);
/* Cells that turned fluid (fresh) or solid (dead) at the last retag, ascending.
 * Writes min(count, cap) entries; *count gets the full count. */
int lbm_fresh_cells(lbm_sim *sim, int32_t *out, int cap, int *count);
int lbm_dead_cells(lbm_sim *sim, int32_t *out, int cap, int *count);

/* ------------------------------------------------------------------ sphere */
/* Host implementations of class RigidSphere (src/solid.js), compiled from the
 * same source as the device integrator. */

/* new RigidSphere(nx, ny, nz, opts) (src/solid.js:9-37).  Pass NaN for an
 * option to take the reference default (cx = nx/2, cy = 0.3 ny, cz = nz/2,
 * r = 0.1 min(nx,ny,nz), density 1.4). */
void lbm_sphere_init(lbm_sphere *s, int nx, int ny, int nz, double cx, double cy, double cz,
                     double r, double density);
/* RigidSphere.recomputeMass (src/solid.js:39-44); setRadius/setDensity assign
 * the field and call it. */
void lbm_sphere_recompute_mass(lbm_sphere *s);
/* RigidSphere.contains (src/solid.js:50-55). */
int lbm_sphere_contains(const lbm_sphere *s, int x, int y, int z);
/* RigidSphere.surfaceVelocity (src/solid.js:57-66). */
void lbm_sphere_surface_velocity(const lbm_sphere *s, int x, int y, int z, double out[3]);
/* RigidSphere.applyImpulseAtCell (src/solid.js:70-80). */
void lbm_sphere_apply_impulse_at_cell(lbm_sphere *s, int x, int y, int z, double px,
                                      double py, double pz);
/* RigidSphere.step(dt, gravity) (src/solid.js:104-210). */
void lbm_sphere_step(lbm_sphere *s, double dt, double gravity);

/* ------------------------------------------------------------------ tables */

/* The lattice tables (src/lattices.js:16-99) as the library uses them. */
int lbm_lattice_tables(int32_t ex27[27], int32_t ey27[27], int32_t ez27[27], float w27[27],
                       int32_t opp27[27], int32_t ex7[7], int32_t ey7[7], int32_t ez7[7],
                       float w7[7], int32_t opp7[7]);

/* ------------------------------------------------------------------ timing */

/* The handle's HIP stream (a hipStream_t), for callers that time work with
 * HIP events on it. */
void *lbm_stream(lbm_sim *sim);
/* Blocks until the handle's queued work is done. */
int lbm_synchronize(lbm_sim *sim);
/* When on, lbm_step records HIP events around each launch of its kernels and
 * accumulates per-kernel times, read with lbm_kernel_time. */
int lbm_set_timing(lbm_sim *sim, int on);
/* Total milliseconds and launch count of a kernel since timing was turned on.
 * Names: "stream_collide", "phase_stream", "band", "body", "collide_all",
 * "stream_macro". */
int lbm_kernel_time(lbm_sim *sim, const char *name, double *total_ms, int64_t *launches);

/* Launch graphs of lbm_step.  No reference counterpart (the reference calls
 * its passes from JavaScript, src/sim.js:202-210).  Every step of a call but
 * the first and the last launches the same kernels with the same arguments
 * except the step parity; runs of `steps` such steps (even, >= 2) are captured
 * once into a HIP graph and replayed, here and in later calls with identical
 * launch arguments.  steps = 0 launches every kernel on its own.  Results are
 * the same either way, bit for bit.  Timing (lbm_set_timing) launches kernels
 * on their own.  Default 2. */
int lbm_set_graph_steps(lbm_sim *sim, int steps);
/* Graph captures and graph replays since the handle was created. */
int lbm_graph_counts(const lbm_sim *sim, int64_t *captures, int64_t *replays);
/* Cells per lane of the fused stream kernel with 16-bit storage: 2 (default)
 * reads and writes each population plane in aligned 4-byte pairs of cells, 1
 * in single 2-byte elements.  No reference counterpart.  Both give the same
 * results bit for bit at every fluid cell (populations of non-fluid cells are
 * unspecified); float32 storage always runs one cell per lane. */
int lbm_set_stream_lanes(lbm_sim *sim, int cells_per_lane);

/* Launch shape of the fused stream kernel for one simulation: 1 launches the
 * tiles away from the sphere by a kernel that carries no tag reads and no
 * momentum exchange, and the tiles of the sphere's band rows by the full
 * kernel; 0 launches every tile with the full kernel; 2 (default; float32
 * storage with rows of at most 256 cells, otherwise as 1) runs the row march:
 * one launch computes phi(t+1) and streams and collides every cell outside the
 * sphere's band box, row by row along z, beside a launch over the band box's
 * cells.  No reference counterpart; results are the same bit for bit. */
int lbm_set_stream_split(lbm_sim *sim, int mode);
/* Workgroup shape of the row march (mode 2 above): each workgroup marches
 * `rows` rows (1 or 2; default 2) through nz / segments planes; segments 0
 * (default) chooses segments of about eight planes, and at least 4096 / rows
 * waves in all.  No reference counterpart; results do not depend on it. */
int lbm_set_march_shape(lbm_sim *sim, int rows, int segments);

/* Measured device bandwidth of a plain copy: copies `bytes` (rounded down to 16)
 * between two device buffers `reps` times with one float4 per thread and
 * returns the read+write rate in GB/s in *gbps.  No reference counterpart: it
 * is the measured ceiling the roofline figures are compared with, and the
 * kernel of known size that calibrates hardware traffic counters. */
int lbm_probe_copy_bandwidth(int64_t bytes, int reps, double *gbps);
/* Measured rate of the fused stream kernels' population pattern alone: the
 * 27 + 7 planes of an nx x ny x nz grid pulled from their shifted sources and
 * stored, in the kernels' tiles, order and cache policy (every pull of a lane
 * issued before its first store), with no arithmetic (storage: LBM_STORAGE_F32 one cell per lane, or a 16-bit form two cells per
 * lane); `reps` timed passes after one warm-up, read+write rate in GB/s over
 * 34 planes x element bytes x cells.  No reference counterpart: the ceiling the
 * stream kernels' memory pattern allows, beside the roofline. */
int lbm_probe_pull_bandwidth(int nx, int ny, int nz, int storage, int reps, double *gbps);

/* Library version string. */
const char *lbm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
