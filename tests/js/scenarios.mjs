// Counterparts of the reference's scenario scripts (test/smoke.mjs,
// test/profile.mjs, test/repro_presets.mjs, test/drag_stress.mjs,
// test/settle.mjs) through the drop-in modules.  Where the reference only
// prints, these assert what its output is read for: finite fields and bodies,
// maxU <= 1, phi in [-0.01, 1.01].  Exit status 1 on any failure.
import { Sim } from '../../claude-vibe-code-raytracing_amd/node/sim.js';
import { RigidSphere } from '../../claude-vibe-code-raytracing_amd/node/solid.js';

let failures = 0;
function check(cond, msg) {
  if (!cond) { failures++; console.log('FAIL ' + msg); }
}

function scan(sim) {
  let nan = 0, maxU = 0, lo = Infinity, hi = -Infinity;
  const { phi, rho, ux, uy, uz } = sim;
  for (let i = 0; i < sim.n; i++) {
    if (!Number.isFinite(phi[i]) || !Number.isFinite(rho[i]) || !Number.isFinite(ux[i]) ||
        !Number.isFinite(uy[i]) || !Number.isFinite(uz[i])) nan++;
    const u = Math.hypot(ux[i], uy[i], uz[i]);
    if (u > maxU) maxU = u;
    if (phi[i] < lo) lo = phi[i];
    if (phi[i] > hi) hi = phi[i];
  }
  return { nan, maxU, lo, hi };
}

function bodyOk(s) {
  return Number.isFinite(s.cx) && Number.isFinite(s.cy) && Number.isFinite(s.cz);
}

// smoke.mjs: 12^3, 80 coupled steps, pass by pass.
{
  const sim = new Sim(12, 12, 12);
  sim.initFlat(6);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.0012; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(12, 12, 12, { cx: 6, cy: 3, cz: 6, r: 2.0, density: 1.5 });
  for (let s = 0; s < 80; s++) {
    sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity);
    const r = scan(sim);
    check(r.nan === 0, `smoke: NaN at step ${s}`);
    check(r.lo >= -0.01 && r.hi <= 1.01, `smoke: phi out of range at step ${s}`);
  }
  check(sim.phi[sim.idx(6, 10, 6)] > 0.8 && sim.phi[sim.idx(6, 1, 6)] < 0.1,
    'smoke: liquid stays at the bottom');
  console.log(`smoke: solid (${solid.cx.toFixed(2)}, ${solid.cy.toFixed(2)}, ${solid.cz.toFixed(2)})`);
  sim.dispose();
}

// profile.mjs: no solid, 60 steps; the mean phi is conserved closely.
{
  const sim = new Sim(8, 12, 8);
  sim.initFlat(6);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.0012; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  let avg0 = 0;
  for (let i = 0; i < sim.n; i++) avg0 += sim.phi[i];
  sim.run(null, 60);
  let avg = 0;
  for (let i = 0; i < sim.n; i++) avg += sim.phi[i];
  check(scan(sim).nan === 0, 'profile: NaN');
  check(Math.abs(avg - avg0) / avg0 < 0.02, `profile: phi mass drift ${avg0} -> ${avg}`);
  sim.dispose();
}

function presetScene(name, setup, steps, fused) {
  const NX = 40, NY = 28, NZ = 40;
  const sim = new Sim(NX, NY, NZ);
  sim.tau = 0.85; sim.tauPhi = 0.7; sim.sigma = 0.008;
  sim.gravity = 0.0015; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(NX, NY, NZ, { cx: NX * 0.5, cy: NY * 0.2, cz: NZ * 0.5, r: 4, density: 1.6 });
  setup(sim, solid, NX, NY, NZ);
  let peak = 0;
  for (let s = 1; s <= steps; s += fused ? 10 : 1) {
    if (fused) sim.run(solid, 10);
    else { sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity); }
    const r = scan(sim);
    if (r.maxU > peak) peak = r.maxU;
    check(r.nan === 0 && bodyOk(solid), `${name}: blew up at step ${s}`);
    check(r.maxU <= 1.0, `${name}: maxU ${r.maxU} at step ${s}`);
    if (r.nan) break;
  }
  console.log(`${name}: cy=${solid.cy.toFixed(2)} peak maxU=${peak.toFixed(3)}`);
  sim.dispose();
}

// repro_presets.mjs: drop, dam (host writes phi/phiPrev/h + seedHydrostatic), rise.
const drop = (sim, solid, NX, NY, NZ) => {
  sim.initFlat(NY * 0.25);
  solid.cx = NX * 0.5; solid.cy = NY * 0.35; solid.cz = NZ * 0.5;
  solid.setRadius(4.0); solid.setDensity(2.0);
};
const dam = (sim, solid, NX, NY, NZ) => {
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
};
const rise = (sim, solid, NX, NY, NZ) => {
  sim.initFlat(NY * 0.25);
  solid.cx = NX * 0.5; solid.cy = NY * 0.65; solid.cz = NZ * 0.5;
  solid.setRadius(4.0); solid.setDensity(0.5);
};
presetScene('drop', drop, 120, false);
presetScene('dam', dam, 120, false);
presetScene('rise', rise, 120, false);
// settle.mjs: the same presets much longer, on the fused path.
presetScene('settle-drop', drop, 1500, true);
presetScene('settle-rise', rise, 1500, true);

// drag_stress.mjs: drag through the interface and back (700 steps, fused path).
{
  const NX = 40, NY = 28, NZ = 40;
  const sim = new Sim(NX, NY, NZ);
  sim.tau = 0.8; sim.tauPhi = 0.7; sim.sigma = 0.008;
  sim.gravity = 0.0015; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  sim.initFlat(NY * 0.25);
  const solid = new RigidSphere(NX, NY, NZ, { cx: NX * 0.5, cy: NY * 0.65, cz: NZ * 0.5, r: 4, density: 1.4 });
  solid.dragging = true;
  const path = [
    { steps: 100, x: 20, y: 24, z: 20 }, { steps: 100, x: 30, y: 24, z: 20 },
    { steps: 200, x: 30, y: 4, z: 20 }, { steps: 100, x: 20, y: 4, z: 20 },
    { steps: 200, x: 20, y: 18, z: 20 },
  ];
  for (const leg of path) {
    solid.dragTargetX = leg.x; solid.dragTargetY = leg.y; solid.dragTargetZ = leg.z;
    for (let s = 0; s < leg.steps; s += 20) {
      sim.run(solid, 20);
      check(Number.isFinite(solid.cy) && Number.isFinite(solid.vy), 'drag: body blew up');
      const r = scan(sim);
      check(r.maxU <= 1.0 && r.nan === 0, `drag: maxU ${r.maxU}`);
    }
  }
  console.log(`drag: ended at (${solid.cx.toFixed(2)}, ${solid.cy.toFixed(2)}, ${solid.cz.toFixed(2)})`);
  sim.dispose();
}

console.log(failures ? `FAILED ${failures}` : 'ALL SCENARIOS OK');
process.exit(failures ? 1 : 0);
