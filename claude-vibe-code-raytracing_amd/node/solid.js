// RigidSphere with the surface of the reference's src/solid.js.  The state is a
// Float64Array(27) laid out as struct lbm_sphere, so the native library reads
// and updates it in place; every method runs the library's host implementation
// of the reference's arithmetic.
import { native } from './native.js';

const FIELDS = ['nx', 'ny', 'nz', 'cx', 'cy', 'cz', 'r', 'vx', 'vy', 'vz', 'wx', 'wy', 'wz',
  'density', 'volume', 'mass', 'I', 'dragging', 'dragTargetX', 'dragTargetY', 'dragTargetZ',
  'Fx', 'Fy', 'Fz', 'Tx', 'Ty', 'Tz'];

function opt(o, k) {
  // `opts.k ?? NaN` (src/solid.js:11-14, 24); NaN asks the library for the default.
  const v = o[k];
  return v === undefined || v === null ? NaN : v;
}

export class RigidSphere {
  // new RigidSphere(nx, ny, nz, opts) (src/solid.js:9-37)
  constructor(nx, ny, nz, opts = {}) {
    this.state = new Float64Array(FIELDS.length);
    this._busy = false;  // stepped by a Sim.runAsync that has not settled
    native.sphereInit(this.state, nx, ny, nz, opt(opts, 'cx'), opt(opts, 'cy'), opt(opts, 'cz'),
      opt(opts, 'r'), opt(opts, 'density'));
  }

  // Writes throw while a Sim.runAsync steps this sphere on a worker thread.
  _idle() {
    if (this._busy) throw new Error('RigidSphere is in use by Sim.runAsync until its promise settles');
  }

  // src/solid.js:39-47
  recomputeMass() { this._idle(); native.sphereRecomputeMass(this.state); }
  setRadius(r) { this.r = r; this.recomputeMass(); }
  setDensity(d) { this.density = d; this.recomputeMass(); }

  // src/solid.js:50-80
  contains(x, y, z) { return native.sphereContains(this.state, x, y, z); }
  surfaceVelocity(x, y, z, out) { return native.sphereSurfaceVelocity(this.state, x, y, z, out); }
  applyImpulseAtCell(x, y, z, px, py, pz) {
    this._idle();
    native.sphereApplyImpulse(this.state, x, y, z, px, py, pz);
  }

  // src/solid.js:87-101: dead cells of sim's last retag, in index order.
  absorbDeadCells(sim) { this._idle(); sim._absorbDead(this); }

  // src/solid.js:104-210
  step(dt, gravity) { this._idle(); native.sphereStep(this.state, dt, gravity); }
}

FIELDS.forEach((k, idx) => {
  Object.defineProperty(RigidSphere.prototype, k, {
    get() { return k === 'dragging' ? this.state[idx] !== 0 : this.state[idx]; },
    set(v) {
      this._idle();
      this.state[idx] = k === 'dragging' ? (v ? 1 : 0) : v;
    },
    enumerable: true,
  });
});
