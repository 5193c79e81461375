// Sim.runAsync (lbm_step on a worker thread, a Promise) gives the same state as
// Sim.run, and the event loop keeps running while it steps; until it settles the
// Sim and its sphere refuse every use; checkFields reports a finite state and
// |u| under the reference's cap.
import { Sim } from '../../claude-vibe-code-raytracing_amd/node/sim.js';
import { RigidSphere } from '../../claude-vibe-code-raytracing_amd/node/solid.js';

function fail(m) { console.error('FAIL', m); process.exit(1); }

function make() {
  const sim = new Sim(40, 28, 40);
  sim.tau = 0.6; sim.tauPhi = 0.58; sim.sigma = 0.01; sim.gravity = 1e-5;
  sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  sim.initFlat(14);
  const solid = new RigidSphere(40, 28, 40, { cx: 20, cy: 9, cz: 20, r: 5, density: 2.0 });
  return { sim, solid };
}

async function main() {
  const a = make(), b = make();
  a.sim.run(a.solid, 60);
  let ticks = 0;
  const timer = setInterval(() => { ticks++; }, 0);
  const pending = b.sim.runAsync(b.solid, 30);
  // While the worker thread steps them, the Sim and its sphere refuse use.
  let refused = 0;
  for (const use of [() => b.sim.phi, () => b.sim.run(b.solid, 1), () => b.sim.runAsync(b.solid, 1),
    () => b.sim.packVolume(new Uint8Array(2 * b.sim.n)), () => { b.solid.vy = 0; },
    () => b.solid.step(1.0, 1e-5)]) {
    try { use(); } catch (e) { refused++; }
  }
  if (refused !== 6) fail(`busy guard refused ${refused} of 6 uses`);
  await pending;
  await b.sim.runAsync(b.solid, 30);
  clearInterval(timer);
  for (const k of ['phi', 'rho', 'ux', 'uy', 'uz', 'tag']) {
    const x = a.sim[k], y = b.sim[k];
    for (let i = 0; i < x.length; i++) if (!Object.is(x[i], y[i]) && !(x[i] === y[i])) fail(`${k}[${i}]`);
  }
  if (a.solid.cy !== b.solid.cy || a.solid.vy !== b.solid.vy) fail('sphere');
  const c = b.sim.checkFields();
  if (!c.finite || !(c.maxU > 0 && c.maxU <= 0.2 + 1e-7)) fail(`checkFields ${JSON.stringify(c)}`);
  console.log(`ASYNC OK ticks=${ticks} maxU=${c.maxU}`);
  a.sim.dispose();
  b.sim.dispose();
}

main().catch((e) => fail(e.stack || String(e)));
