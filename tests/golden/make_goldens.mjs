// Writes the golden vectors in tests/golden/ by running the reference simulator
// (read-only, from /root/reference) in this container.  Run with Node 12:
//
//   node --harmony-nullish --experimental-loader ./tests/golden/loader.mjs \
//        tests/golden/make_goldens.mjs
//
// (--harmony-nullish enables the `??` operator that src/solid.js uses.)
//
// Each case writes <case>.json (parameters, per-step sphere state, per-step
// fresh/dead cell lists, and a table of contents) and <case>.bin.gz (the
// arrays named in the table of contents, concatenated, little-endian).
import fs from 'fs';
import zlib from 'zlib';
import path from 'path';
import { Sim } from '/root/reference/src/sim.js';
import { RigidSphere } from '/root/reference/src/solid.js';
import * as L from '/root/reference/src/lattices.js';

const OUT = path.dirname(new URL(import.meta.url).pathname);

const SPHERE_FIELDS = ['nx', 'ny', 'nz', 'cx', 'cy', 'cz', 'r', 'vx', 'vy', 'vz',
  'wx', 'wy', 'wz', 'density', 'volume', 'mass', 'I', 'dragging',
  'dragTargetX', 'dragTargetY', 'dragTargetZ', 'Fx', 'Fy', 'Fz', 'Tx', 'Ty', 'Tz'];

function sphereState(s) {
  return SPHERE_FIELDS.map((k) => (k === 'dragging' ? (s[k] ? 1 : 0) : s[k]));
}

class Writer {
  constructor(name) { this.name = name; this.toc = []; this.chunks = []; this.off = 0; this.meta = {}; }
  add(key, arr) {
    const bytes = Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength);
    const dtype = arr instanceof Float32Array ? 'f4' : arr instanceof Float64Array ? 'f8'
      : arr instanceof Uint8Array ? 'u1' : arr instanceof Int32Array ? 'i4' : null;
    if (!dtype) throw new Error('dtype ' + key);
    this.toc.push({ key, dtype, count: arr.length, offset: this.off });
    this.chunks.push(Buffer.from(bytes));
    this.off += bytes.length;
  }
  addSim(prefix, sim, fields) {
    for (const f of fields) this.add(`${prefix}/${f}`, sim[f]);
  }
  write() {
    const blob = zlib.gzipSync(Buffer.concat(this.chunks), { level: 9 });
    fs.writeFileSync(path.join(OUT, this.name + '.bin.gz'), blob);
    this.meta.toc = this.toc;
    fs.writeFileSync(path.join(OUT, this.name + '.json'), JSON.stringify(this.meta));
    console.log(`${this.name}: ${blob.length} bytes gz, ${this.toc.length} arrays`);
  }
}

function lists(sim) {
  const fresh = [], dead = [];
  const { nx, ny, nz } = sim;
  for (let z = 1; z < nz - 1; z++)
    for (let y = 1; y < ny - 1; y++)
      for (let x = 1; x < nx - 1; x++) {
        const i = ((z * ny) + y) * nx + x;
        if (sim.tag[i] === L.TAG_FLUID && sim.tagPrev[i] === L.TAG_SOLID) fresh.push(i);
        if (sim.tag[i] === L.TAG_SOLID && sim.tagPrev[i] === L.TAG_FLUID) dead.push(i);
      }
  return { fresh, dead };
}

function params(sim) {
  const o = {};
  for (const k of ['tau', 'tauPhi', 'sigma', 'gravity', 'rhoL', 'rhoG', 'W', 'mobility']) o[k] = sim[k];
  return o;
}

const MACROS = ['phi', 'phiPrev', 'rho', 'ux', 'uy', 'uz', 'tag', 'tagPrev'];
const FULL = ['f', 'h', ...MACROS, 'usx', 'usy', 'usz', 'gx', 'gy', 'gz', 'lap'];

// ---------------------------------------------------------------- tables
{
  const w = new Writer('lattice_tables');
  w.add('W27', L.W27); w.add('W7', L.W7);
  w.add('EX27', new Int32Array(L.EX27)); w.add('EY27', new Int32Array(L.EY27));
  w.add('EZ27', new Int32Array(L.EZ27)); w.add('OPP27', new Int32Array(L.OPP27));
  w.add('EX7', new Int32Array(L.EX7)); w.add('EY7', new Int32Array(L.EY7));
  w.add('EZ7', new Int32Array(L.EZ7)); w.add('OPP7', new Int32Array(L.OPP7));
  w.meta.consts = { CS2_27: L.CS2_27, INV_CS2_27: L.INV_CS2_27, INV_2CS4_27: L.INV_2CS4_27,
    CS2_7: L.CS2_7, INV_CS2_7: L.INV_CS2_7 };
  w.write();
}

// Runs `steps` coupled steps in the driver order of src/worker.js:96-100,
// recording the sphere state and fresh/dead lists after every step and the
// requested fields at the requested steps.
function trajectory(w, sim, solid, steps, snapAt, fields) {
  w.meta.params = params(sim);
  w.meta.sphere0 = sphereState(solid);
  w.meta.sphere_fields = SPHERE_FIELDS;
  w.addSim('step0', sim, fields);
  const sph = [], fresh = [], dead = [], forces = [];
  for (let s = 1; s <= steps; s++) {
    sim.step(solid);
    const l = lists(sim);
    fresh.push(l.fresh); dead.push(l.dead);
    // Force/torque from the link impulses alone, before dead cells and gravity.
    const fLinks = [solid.Fx, solid.Fy, solid.Fz, solid.Tx, solid.Ty, solid.Tz];
    solid.absorbDeadCells(sim);
    const fAll = [solid.Fx, solid.Fy, solid.Fz, solid.Tx, solid.Ty, solid.Tz];
    forces.push([fLinks, fAll]);
    solid.step(1.0, sim.gravity);
    sph.push(sphereState(solid));
    if (snapAt.includes(s)) w.addSim(`step${s}`, sim, fields);
  }
  w.meta.sphere = sph; w.meta.fresh = fresh; w.meta.dead = dead; w.meta.forces = forces;
  w.meta.snap_steps = [0, ...snapAt];
}

// ---------------------------------------------------------------- smoke 12^3
// test/smoke.mjs:7-19 -- note initFlat runs before the parameters are set.
{
  const w = new Writer('smoke12');
  const sim = new Sim(12, 12, 12);
  sim.initFlat(6);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.0012; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(12, 12, 12, { cx: 6, cy: 3, cz: 6, r: 2.0, density: 1.5 });
  w.meta.setup = 'test/smoke.mjs: Sim(12,12,12); initFlat(6) with default params; then params; sphere (6,3,6) r=2 density 1.5';
  trajectory(w, sim, solid, 80, [1, 2, 10, 80], FULL);
  w.write();
}

// ---------------------------------------------------------------- per-pass 12^3
// Smoke scene advanced 10 coupled steps, then step 11 pass by pass.  Records the
// full state before each pass and the state after it.
{
  const w = new Writer('passes12');
  const sim = new Sim(12, 12, 12);
  sim.initFlat(6);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.0012; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(12, 12, 12, { cx: 6, cy: 3, cz: 6, r: 2.0, density: 1.5 });
  for (let s = 0; s < 10; s++) { sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity); }
  // Give the sphere a sideways kick and spin so step 11 has fresh and dead
  // cells and a moving wall.
  solid.vx = 0.05; solid.vz = -0.03; solid.wy = 0.01; solid.wx = -0.02;
  solid.step(1.0, sim.gravity);
  // A nonzero external body force on a few cells exercises bodyF.
  for (let i = 0; i < sim.n; i += 37) { sim.bodyFx[i] = 1e-4; sim.bodyFy[i] = -2e-4; sim.bodyFz[i] = 5e-5; }
  w.meta.params = params(sim);
  const ALL = ['f', 'f2', 'h', 'h2', ...MACROS, 'usx', 'usy', 'usz', 'gx', 'gy', 'gz', 'lap',
    'bodyFx', 'bodyFy', 'bodyFz'];
  w.meta.sphere_fields = SPHERE_FIELDS;
  w.meta.sphere_in = sphereState(solid);
  w.addSim('in', sim, ALL);
  // Each pass records the arrays it writes (the others are unchanged by it).
  const passes = [
    ['retagCells', () => sim.retagCells(solid), ['tag', 'tagPrev', 'phiPrev', 'usx', 'usy', 'usz']],
    ['handleFreshCells', () => sim.handleFreshCells(solid), ['f', 'h', 'phi', 'phiPrev', 'rho', 'ux', 'uy', 'uz']],
    ['computePhiGradients', () => sim.computePhiGradients(), ['gx', 'gy', 'gz', 'lap']],
    ['collidePhase', () => sim.collidePhase(), ['h2']],
    ['streamPhase', () => sim.streamPhase(), ['h']],
    ['collideHydro', () => sim.collideHydro(), ['f2']],
    ['streamHydroAndMacro', () => sim.streamHydroAndMacro(solid),
      ['f', 'rho', 'phi', 'ux', 'uy', 'uz', 'bodyFx', 'bodyFy', 'bodyFz']],
  ];
  w.meta.passes = passes.map((p) => p[0]);
  w.meta.pass_outputs = Object.fromEntries(passes.map((p) => [p[0], p[2]]));
  for (const [name, fn, outs] of passes) {
    fn();
    w.addSim(name, sim, outs);
  }
  w.meta.sphere_after_stream = sphereState(solid);
  solid.absorbDeadCells(sim);
  w.meta.sphere_after_absorb = sphereState(solid);
  solid.step(1.0, sim.gravity);
  w.meta.sphere_after_step = sphereState(solid);
  const buf = new Uint8Array(sim.n * 2);
  sim.packVolume(buf, null);
  w.add('packVolume', buf);
  w.write();
}

// ---------------------------------------------------------------- config 0
// 40x28x40, parameters set BEFORE initFlat(14): tau .6, tauPhi .58 (mobility
// 0.02 = cs7^2 (tauPhi - 1/2)), sigma .01, g 1e-5, rhoL 1, rhoG .1, W 4.
// Sphere r=5, density 2, at (20,22,20), at rest.  200 steps.
{
  const w = new Writer('config0');
  const sim = new Sim(40, 28, 40);
  sim.tau = 0.6; sim.tauPhi = 0.58; sim.sigma = 0.01; sim.gravity = 1e-5;
  sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  sim.initFlat(14);
  const solid = new RigidSphere(40, 28, 40, { cx: 20, cy: 22, cz: 20, r: 5, density: 2.0 });
  w.meta.setup = 'config 0: params first, initFlat(14), sphere (20,22,20) r=5 density 2';
  trajectory(w, sim, solid, 200, [1, 10, 200], ['phi', 'rho', 'ux', 'uy', 'uz', 'tag']);
  w.write();
}

// ---------------------------------------------------------------- dam preset
// src/worker.js:54-79 / test/repro_presets.mjs:54-71: host writes phi, phiPrev
// and h directly, then seedHydrostatic().  Smaller box to keep the fixture small.
{
  const w = new Writer('dam16');
  const NX = 16, NY = 12, NZ = 14;
  const sim = new Sim(NX, NY, NZ);
  sim.tau = 0.85; sim.tauPhi = 0.7; sim.sigma = 0.008;
  sim.gravity = 0.0015; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(NX, NY, NZ, { cx: NX * 0.5, cy: NY * 0.2, cz: NZ * 0.5, r: 4, density: 1.6 });
  sim.initFlat(NY + 1);
  const w7 = [1 / 4, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8];
  for (let z = 1; z < NZ - 1; z++)
    for (let y = 1; y < NY - 1; y++)
      for (let x = 1; x < NX - 1; x++) {
        const i = sim.idx(x, y, z);
        const p = (x < NX * 0.35 && y > NY * 0.35) ? 1 : 0;
        sim.phi[i] = p; sim.phiPrev[i] = p;
        for (let k = 0; k < 7; k++) sim.h[i * 7 + k] = w7[k] * p;
      }
  sim.seedHydrostatic();
  solid.cx = NX * 0.7; solid.cy = NY * 0.6; solid.cz = NZ * 0.5;
  solid.setRadius(3.5); solid.setDensity(0.5);
  w.meta.setup = 'dam preset on 16x12x14 (worker.js:54-79 pattern)';
  trajectory(w, sim, solid, 40, [1, 40], ['f', 'h', ...MACROS]);
  w.write();
}

// ---------------------------------------------------------------- no solid
// test/profile.mjs: 8x12x8, sim.step(null), 60 steps.
{
  const w = new Writer('profile8');
  const sim = new Sim(8, 12, 8);
  sim.initFlat(6);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.0012; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  w.meta.params = params(sim);
  w.addSim('step0', sim, ['f', 'h', ...MACROS]);
  for (let s = 1; s <= 60; s++) sim.step(null);
  w.addSim('step60', sim, ['f', 'h', ...MACROS]);
  w.meta.snap_steps = [0, 60];
  w.write();
}

// ---------------------------------------------------------------- drag spring
// test/drag_stress.mjs pattern on a 20x16x20 box: dragging through the interface
// exercises the spring force, the acceleration cap and the speed cap.
{
  const w = new Writer('drag20');
  const NX = 20, NY = 16, NZ = 20;
  const sim = new Sim(NX, NY, NZ);
  sim.tau = 0.8; sim.tauPhi = 0.7; sim.sigma = 0.008;
  sim.gravity = 0.0015; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  sim.initFlat(NY * 0.25);
  const solid = new RigidSphere(NX, NY, NZ, { cx: NX * 0.5, cy: NY * 0.65, cz: NZ * 0.5, r: 3, density: 1.4 });
  solid.dragging = true;
  solid.dragTargetX = 14; solid.dragTargetY = 3; solid.dragTargetZ = 9;
  w.meta.setup = 'drag spring toward (14,3,9) on 20x16x20';
  trajectory(w, sim, solid, 60, [60], ['phi', 'rho', 'ux', 'uy', 'uz', 'tag']);
  w.write();
}
