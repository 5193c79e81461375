// RG8 frame delivery: frame_server.js + remote_worker.js end to end.  The page
// side runs remote_worker.js with a browser-like WebSocket over ws.js; every
// frame it posts must equal packVolume of the same scene stepped locally to the
// same step count (RG8 bytes and the solid object), pause must stop the steps,
// and a parameter message must reach the simulation.
import { createFrameServer, makeScene, applyPreset } from '../../claude-vibe-code-raytracing_amd/node/frame_server.js';
import { connect } from '../../claude-vibe-code-raytracing_amd/node/ws.js';

class BrowserWebSocket {
  constructor(url) {
    this.onopen = null;
    this.onmessage = null;
    this.onclose = null;
    connect(url).then((c) => {
      this.c = c;
      c.on('message', (d, bin) => {
        if (this.onmessage) {
          this.onmessage({ data: bin ? d.buffer.slice(d.byteOffset, d.byteOffset + d.byteLength) : d });
        }
      });
      c.on('close', () => { if (this.onclose) this.onclose(); });
      if (this.onopen) this.onopen();
    });
  }
  send(m) { this.c.send(m); }
  close() { this.c.close(); }
}

function fail(msg) { console.error('FAIL', msg); process.exit(1); }

async function main() {
  const srv = await createFrameServer({ port: 0 });
  const frames = [];
  let waiter = null;
  globalThis.self = {
    location: { search: `?server=ws://127.0.0.1:${srv.port}` },
    WebSocket: BrowserWebSocket,
    postMessage: (m) => { frames.push(m); if (waiter) waiter(); },
  };
  await import('../../claude-vibe-code-raytracing_amd/node/remote_worker.js');
  const waitFrames = (k) => new Promise((res) => {
    const check = () => { if (frames.length >= k) { waiter = null; res(); } };
    waiter = check;
    check();
  });
  const NX = 40, NY = 28, NZ = 40;
  self.onmessage({ data: { type: 'init', nx: NX, ny: NY, nz: NZ, preset: 'drop', substeps: 2 } });
  await waitFrames(6);
  // Local replay of the same scene in ticks of 2 steps.
  const { sim, solid } = makeScene(NX, NY, NZ);
  applyPreset(sim, solid, 'drop');
  const vol = new Uint8Array(2 * NX * NY * NZ);
  let steps = 0;
  const seen = frames.slice(0, 6);
  for (const f of seen) {
    if (f.type !== 'frame' || f.vol.length !== vol.length) fail('frame shape');
    while (steps < f.steps) { sim.run(solid, 2); steps += 2; }
    sim.packVolume(vol, null);
    for (let i = 0; i < vol.length; i++) if (vol[i] !== f.vol[i]) fail(`RG8 differs at step ${f.steps} byte ${i}`);
    const s = f.solid;
    if (s.cx !== solid.cx || s.cy !== solid.cy || s.cz !== solid.cz || s.vy !== solid.vy || s.r !== solid.r) {
      fail(`solid differs at step ${f.steps}`);
    }
  }
  if (seen[seen.length - 1].steps < 10) fail('too few steps');
  // Hand a buffer back, as main.js does.
  const back = frames[0].vol;
  self.onmessage({ data: { type: 'releaseBuf', buffer: back.buffer } });
  // pause: the step count stops
  self.onmessage({ data: { type: 'pause' } });
  const k0 = frames.length;
  await waitFrames(k0 + 4);
  const a = frames[frames.length - 2].steps, b = frames[frames.length - 1].steps;
  if (a !== b) fail('steps advance while paused');
  // param + resume
  self.onmessage({ data: { type: 'param', name: 'gravity', value: 0.003 } });
  self.onmessage({ data: { type: 'resume' } });
  const k1 = frames.length;
  await waitFrames(k1 + 4);
  if (!(frames[frames.length - 1].steps > b)) fail('steps do not resume');
  console.log(`OK frames=${frames.length} last_steps=${frames[frames.length - 1].steps}`);
  sim.dispose();
  await srv.close();
  process.exit(0);
}

main().catch((e) => fail(e && e.stack ? e.stack : String(e)));
