// Counterpart of test/bench.mjs through the drop-in: same scene and step
// counts, timed per coupled step both pass by pass (sim.step + absorbDeadCells +
// solid.step) and on the fused path (sim.run).  Prints one JSON line.
import { performance } from 'perf_hooks';
import { Sim } from '../../claude-vibe-code-raytracing_amd/node/sim.js';
import { RigidSphere } from '../../claude-vibe-code-raytracing_amd/node/solid.js';

function scene() {
  const sim = new Sim(40, 28, 40);
  sim.initFlat(14);
  sim.tau = 0.7; sim.tauPhi = 0.7; sim.sigma = 0.012;
  sim.gravity = 0.00135; sim.rhoL = 1.0; sim.rhoG = 0.1; sim.W = 4.0;
  const solid = new RigidSphere(40, 28, 40, { cx: 20, cy: 6, cz: 20, r: 4, density: 1.6 });
  return { sim, solid };
}
const N = 40 * 28 * 40;
const out = {};
{
  const { sim, solid } = scene();
  for (let i = 0; i < 20; i++) { sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity); }
  const t0 = performance.now();
  for (let i = 0; i < 80; i++) { sim.step(solid); solid.absorbDeadCells(sim); solid.step(1.0, sim.gravity); }
  const dt = (performance.now() - t0) / 80;
  out.passes_ms_per_step = dt; out.passes_mlups = N / dt * 1e-3;
  sim.dispose();
}
{
  const { sim, solid } = scene();
  sim.run(solid, 20);
  const t0 = performance.now();
  sim.run(solid, 80);
  const dt = (performance.now() - t0) / 80;
  out.fused_ms_per_step = dt; out.fused_mlups = N / dt * 1e-3;
  sim.dispose();
}
console.log(JSON.stringify(out));
