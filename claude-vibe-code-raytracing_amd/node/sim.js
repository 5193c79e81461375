// Sim with the surface of the reference's src/sim.js, running on liblbm (HIP).
//
// Parameters are plain fields read at every pass (src/sim.js:88-95).  The typed
// arrays (phi, rho, ux, ..., f, h) are host mirrors of the device state with a
// stable identity per field:
//   * a field is downloaded the first time it is read and, from then on, kept
//     current after every native call (a "watched" field);
//   * before every native call, each watched mirror is compared with the copy
//     last exchanged with the device and uploaded if the caller wrote to it
//     (src/worker.js:61-67 writes phi, phiPrev and h, then seedHydrostatic()).
// Populations f and h use the reference's layout f[i*27 + k].
//
// Solids: a RigidSphere from ./solid.js runs on the device.  Any other object
// with contains / surfaceVelocity / applyImpulseAtCell (src/solid.js:4-6) is
// evaluated on the host for the mask, and its momentum exchange is replayed
// through applyImpulseAtCell in the reference's order.
import { native } from './native.js';
import { RigidSphere } from './solid.js';
import { Q27, Q7, EX27, EY27, EZ27, EX7, EY7, EZ7, TAG_FLUID, TAG_SOLID } from './lattices.js';

const FIELD_ID = {
  f: 0, f2: 1, h: 2, h2: 3, rho: 4, ux: 5, uy: 6, uz: 7, phi: 8, phiPrev: 9, tag: 10,
  tagPrev: 11, usx: 12, usy: 13, usz: 14, gx: 15, gy: 16, gz: 17, lap: 18, bodyFx: 19,
  bodyFy: 20, bodyFz: 21,
};
const Q = { f: Q27, f2: Q27, h: Q7, h2: Q7 };
const PARAMS = ['tau', 'tauPhi', 'sigma', 'gravity', 'rhoL', 'rhoG', 'W', 'mobility'];

function sameBytes(a, b) {
  if (a.length !== b.length) return false;
  const x = new Uint32Array(a.buffer, a.byteOffset, a.byteLength >> 2);
  const y = new Uint32Array(b.buffer, b.byteOffset, b.byteLength >> 2);
  for (let i = 0; i < x.length; i++) if (x[i] !== y[i]) return false;
  for (let i = x.length << 2; i < a.byteLength; i++) {
    if (new Uint8Array(a.buffer, a.byteOffset)[i] !== new Uint8Array(b.buffer, b.byteOffset)[i]) return false;
  }
  return true;
}

export class Sim {
  // new Sim(nx, ny, nz) (src/sim.js:30-96)
  constructor(nx, ny, nz) {
    this.nx = nx; this.ny = ny; this.nz = nz;
    const n = nx * ny * nz;
    this.n = n;
    this._h = native.create(nx, ny, nz);
    this._mirror = {};   // field -> typed array handed to the caller
    this._shadow = {};   // field -> copy last exchanged with the device
    this._params = new Float64Array(8);
    this._busy = false;  // a runAsync call has not resolved yet

    const nxny = nx * ny;
    this.neighOffset27 = new Int32Array(Q27);
    this.fOffset27 = new Int32Array(Q27);
    for (let k = 0; k < Q27; k++) {
      const off = -(EZ27[k] * nxny + EY27[k] * nx + EX27[k]);
      this.neighOffset27[k] = off;
      this.fOffset27[k] = off * Q27 + k;
    }
    this.neighOffset7 = new Int32Array(Q7);
    this.hOffset7 = new Int32Array(Q7);
    for (let k = 0; k < Q7; k++) {
      const off = -(EZ7[k] * nxny + EY7[k] * nx + EX7[k]);
      this.neighOffset7[k] = off;
      this.hOffset7[k] = off * Q7 + k;
    }

    // src/sim.js:88-95
    this.tau = 0.6;
    this.tauPhi = 0.7;
    this.sigma = 0.012;
    this.gravity = 0.0010;
    this.rhoL = 1.0;
    this.rhoG = 0.05;
    this.W = 4.0;
    this.mobility = 0.02;
  }

  idx(x, y, z) { return ((z * this.ny) + y) * this.nx + x; }
  rhoFromPhi(p) { return this.rhoG + p * (this.rhoL - this.rhoG); }

  // Frees the device state now instead of at garbage collection.
  dispose() {
    this._idle();
    native.destroy(this._h);
  }

  // Every use of the Sim (and of a RigidSphere it steps) throws while a
  // runAsync call is in flight: the worker thread owns the device state then.
  _idle() {
    if (this._busy) throw new Error('Sim is in use by runAsync until its promise settles');
  }

  // ---------------------------------------------------------------- plumbing
  _alloc(name) {
    const len = this.n * (Q[name] || 1);
    return (name === 'tag' || name === 'tagPrev') ? new Uint8Array(len) : new Float32Array(len);
  }

  _field(name) {
    this._idle();
    let m = this._mirror[name];
    if (!m) {
      m = this._alloc(name);
      native.download(this._h, FIELD_ID[name], m);
      this._mirror[name] = m;
      this._shadow[name] = m.slice();
    }
    return m;
  }

  // Before a native call: parameters, then any mirror the caller changed.
  _push() {
    const p = this._params;
    PARAMS.forEach((k, j) => { p[j] = this[k]; });
    native.setParams(this._h, p);
    for (const name of Object.keys(this._mirror)) {
      const m = this._mirror[name];
      const s = this._shadow[name];
      if (!sameBytes(m, s)) {
        native.upload(this._h, FIELD_ID[name], m);
        s.set(m);
      }
    }
  }

  // After a native call: refresh every watched mirror in place.
  _pull() {
    for (const name of Object.keys(this._mirror)) {
      const m = this._mirror[name];
      native.download(this._h, FIELD_ID[name], m);
      this._shadow[name].set(m);
    }
  }

  _call(fn) {
    this._idle();
    this._push();
    const r = fn();
    this._pull();
    return r;
  }

  // ---------------------------------------------------------------- init
  initFlat(waterlineY) { this._call(() => native.initFlat(this._h, waterlineY)); }
  seedHydrostatic() { this._call(() => native.seedHydrostatic(this._h)); }
  // initFlat with one waterline per (x, z) column (yw[z*nx + x]).
  initWaterlineMap(yw) {
    this._call(() => native.initWaterlineMap(this._h, Float64Array.from(yw)));
  }

  // ---------------------------------------------------------------- passes
  computePhiGradients() { this._call(() => native.gradients(this._h)); }
  collidePhase() { this._call(() => native.collidePhase(this._h)); }
  streamPhase() { this._call(() => native.streamPhase(this._h)); }
  collideHydro() { this._call(() => native.collideHydro(this._h)); }
  handleFreshCells(_solid) { this._call(() => native.handleFresh(this._h)); }

  retagCells(solid) {
    if (solid == null || solid instanceof RigidSphere) {
      this._call(() => native.retag(this._h, solid ? solid.state : null));
      return;
    }
    // Generic solid: evaluate contains / surfaceVelocity on the host in the
    // reference's order (src/sim.js:220-243).
    const { nx, ny, nz } = this;
    const mask = new Uint8Array(this.n);
    const us = new Float32Array(3 * this.n);
    const tmp = [0, 0, 0];
    for (let z = 1; z < nz - 1; z++) {
      for (let y = 1; y < ny - 1; y++) {
        for (let x = 1; x < nx - 1; x++) {
          if (solid.contains(x, y, z)) {
            const i = ((z * ny) + y) * nx + x;
            mask[i] = 1;
            solid.surfaceVelocity(x, y, z, tmp);
            us[3 * i] = tmp[0]; us[3 * i + 1] = tmp[1]; us[3 * i + 2] = tmp[2];
          }
        }
      }
    }
    this._call(() => native.retagMask(this._h, mask, us));
  }

  streamHydroAndMacro(solid) {
    if (solid == null || solid instanceof RigidSphere) {
      this._call(() => native.streamHydroMacro(this._h, solid ? solid.state : null));
      return;
    }
    const links = this._call(() => native.streamHydroMacroLinks(this._h));
    for (let j = 0; j < links.length; j += 7) {
      solid.applyImpulseAtCell(links[j], links[j + 1], links[j + 2],
        links[j + 4], links[j + 5], links[j + 6]);
    }
  }

  // Sim.step(solid) (src/sim.js:202-210)
  step(solid) {
    if (solid == null || solid instanceof RigidSphere) {
      this._call(() => native.simStep(this._h, solid ? solid.state : null));
      return;
    }
    this.retagCells(solid);
    this.handleFreshCells(solid);
    this.computePhiGradients();
    this.collidePhase();
    this.streamPhase();
    this.collideHydro();
    this.streamHydroAndMacro(solid);
  }

  // nsteps coupled steps in the driver order of src/worker.js:96-100
  // (sim.step(solid); solid.absorbDeadCells(sim); solid.step(1, sim.gravity)) on
  // the fused device path.  No reference counterpart: the fast path.
  run(solid, nsteps) {
    if (solid != null && !(solid instanceof RigidSphere)) {
      for (let s = 0; s < nsteps; s++) {
        this.step(solid);
        if (typeof solid.absorbDeadCells === 'function') solid.absorbDeadCells(this);
        if (typeof solid.step === 'function') solid.step(1.0, this.gravity);
      }
      return;
    }
    this._call(() => native.step(this._h, solid ? solid.state : null, nsteps));
  }

  // run() on a worker thread: resolves when the steps are done, with the
  // mirrors refreshed.  The Sim must not be used until then.  Generic solids
  // step synchronously (their callbacks run on this thread).
  runAsync(solid, nsteps) {
    this._idle();
    if (solid != null && !(solid instanceof RigidSphere)) {
      this.run(solid, nsteps);
      return Promise.resolve();
    }
    if (solid) solid._idle();
    this._push();
    const p = native.stepAsync(this._h, solid ? solid.state : null, nsteps);
    this._busy = true;
    if (solid) solid._busy = true;
    const done = () => {
      this._busy = false;
      if (solid) solid._busy = false;
    };
    return p.then(() => { done(); this._pull(); }, (e) => { done(); throw e; });
  }

  // Opt-in state check (not on the hot path): { finite, maxU } over fluid cells.
  checkFields() {
    this._idle();
    this._push();
    const [finite, maxU] = native.checkFields(this._h);
    return { finite, maxU };
  }

  // RigidSphere.absorbDeadCells(sim) for a device sphere.
  _absorbDead(solid) { this._call(() => native.absorbDead(this._h, solid.state)); }

  // Cells that turned fluid / solid at the last retag, ascending.
  freshCells() { this._idle(); return native.freshCells(this._h); }
  deadCells() { this._idle(); return native.deadCells(this._h); }

  // Sim.packVolume(out, _) (src/sim.js:586-600)
  packVolume(out, _solidContains) {
    this._idle();
    this._push();
    native.pack(this._h, out);
  }
}

for (const name of Object.keys(FIELD_ID)) {
  Object.defineProperty(Sim.prototype, name, {
    get() { return this._field(name); },
    enumerable: true,
  });
}

export { TAG_FLUID, TAG_SOLID };
