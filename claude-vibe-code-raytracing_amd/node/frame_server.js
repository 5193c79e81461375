// Frame server: the simulation worker's message protocol (src/worker.js:94-165)
// over a WebSocket, with the simulation on the GPU through the drop-in Sim /
// RigidSphere.  A browser page keeps its main.js and Renderer unchanged and
// runs ./remote_worker.js in place of src/worker.js; each connection is one
// simulation, stepped continuously, and every tick sends one frame.
//
// Messages from the page: JSON text, the worker's commands ({type:'init', nx,
// ny, nz, substeps, preset}, 'preset', 'param', 'drag', 'pause', 'resume';
// 'releaseBuf' is handled in the page-side worker).
// Frames to the page: one binary message, little-endian:
//   0  u32 magic 0x46384752 ('RG8F')     4  u32 nx    8  u32 ny    12 u32 nz
//   16 u32 steps done                    20 u32 frame counter
//   24 f64 cx, cy, cz, r, vx, vy, vz (the frame message's solid object)
//   80 u8[2 nx ny nz] the packVolume RG8 volume (R = phi*255, G = solid mask)
//
//   node frame_server.js [--port 8765] [--host 127.0.0.1]
import http from 'http';
import { pathToFileURL } from 'url';
import { Sim } from './sim.js';
import { RigidSphere } from './solid.js';
import { attachServer } from './ws.js';

export const FRAME_MAGIC = 0x46384752;
export const FRAME_HEADER = 80;

// Simulation parameters and presets of the page's scene (src/worker.js:23-91).
export function makeScene(nx, ny, nz) {
  const sim = new Sim(nx, ny, nz);
  sim.tau = 0.85;
  sim.tauPhi = 0.7;
  sim.sigma = 0.008;
  sim.gravity = 0.0015;
  sim.rhoL = 1.0;
  sim.rhoG = 0.1;
  sim.W = 4.0;
  const solid = new RigidSphere(nx, ny, nz, { cx: nx * 0.5, cy: ny * 0.20, cz: nz * 0.5, r: 4.0, density: 1.6 });
  return { sim, solid };
}

function placeSphere(solid, cx, cy, cz, r, density) {
  solid.cx = cx; solid.cy = cy; solid.cz = cz;
  solid.vx = 0; solid.vy = 0; solid.vz = 0;
  solid.wx = 0; solid.wy = 0; solid.wz = 0;
  solid.setRadius(r);
  solid.setDensity(density);
}

export function applyPreset(sim, solid, name) {
  const { nx, ny, nz } = sim;
  if (name === 'dam') {
    // A water column in one corner, re-seeded hydrostatically.
    const w7 = [1 / 4, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8, 1 / 8];
    sim.initFlat(ny + 1);
    const phi = sim.phi, phiPrev = sim.phiPrev, h = sim.h;
    for (let z = 1; z < nz - 1; z++) {
      for (let y = 1; y < ny - 1; y++) {
        for (let x = 1; x < nx - 1; x++) {
          const i = sim.idx(x, y, z);
          const p = (x < nx * 0.35 && y > ny * 0.35) ? 1 : 0;
          phi[i] = p;
          phiPrev[i] = p;
          for (let k = 0; k < 7; k++) h[i * 7 + k] = w7[k] * p;
        }
      }
    }
    sim.seedHydrostatic();
    placeSphere(solid, nx * 0.7, ny * 0.6, nz * 0.5, 3.5, 0.5);
  } else if (name === 'rise') {
    sim.initFlat(ny * 0.25);
    placeSphere(solid, nx * 0.5, ny * 0.65, nz * 0.5, 4.0, 0.5);
  } else if (name === 'drop') {
    sim.initFlat(ny * 0.25);
    placeSphere(solid, nx * 0.5, ny * 0.35, nz * 0.5, 4.0, 2.0);
  }
}

// One connection: the worker's state machine, ticking on the event loop.
class Session {
  constructor(ws) {
    this.ws = ws;
    this.scene = null;
    this.running = true;
    this.substeps = 2;
    this.steps = 0;
    this.frames = 0;
    this.timer = null;
    this.closed = false;
    ws.on('message', (data, isBinary) => { if (!isBinary) this.onMessage(JSON.parse(data)); });
    ws.on('close', () => this.close());
  }

  onMessage(m) {
    switch (m.type) {
      case 'init': {
        if (this.scene) this.scene.sim.dispose();
        this.scene = makeScene(m.nx | 0, m.ny | 0, m.nz | 0);
        this.substeps = m.substeps !== undefined && m.substeps !== null ? m.substeps | 0 : 2;
        this.steps = 0;
        this.frames = 0;
        const n = this.scene.sim.n;
        this.msg = Buffer.alloc(FRAME_HEADER + 2 * n);
        this.vol = new Uint8Array(this.msg.buffer, this.msg.byteOffset + FRAME_HEADER, 2 * n);
        applyPreset(this.scene.sim, this.scene.solid, m.preset !== undefined && m.preset !== null ? m.preset : 'drop');
        this.schedule();
        break;
      }
      case 'preset':
        if (this.scene) applyPreset(this.scene.sim, this.scene.solid, m.name);
        break;
      case 'param': {
        if (!this.scene) break;
        const { sim, solid } = this.scene;
        if (m.name === 'sigma') sim.sigma = m.value;
        else if (m.name === 'gravity') sim.gravity = m.value;
        else if (m.name === 'rhoRatio') { sim.rhoL = 1.0; sim.rhoG = 1.0 / m.value; }
        else if (m.name === 'solidRho') solid.setDensity(m.value);
        else if (m.name === 'substeps') this.substeps = m.value | 0;
        break;
      }
      case 'drag':
        if (!this.scene) break;
        this.scene.solid.dragging = !!m.active;
        if (m.active) {
          this.scene.solid.dragTargetX = m.x;
          this.scene.solid.dragTargetY = m.y;
          this.scene.solid.dragTargetZ = m.z;
        }
        break;
      case 'pause': this.running = false; break;
      case 'resume': this.running = true; break;
      default: break;
    }
  }

  schedule() {
    if (this.timer || this.closed) return;
    this.timer = setImmediate(() => { this.timer = null; this.tick(); this.schedule(); });
  }

  // Step, pack, send (src/worker.js:94-115).  While the socket still holds two
  // unsent frames the frame is skipped (the page only shows the latest).
  tick() {
    if (!this.scene || this.closed) return;
    const { sim, solid } = this.scene;
    if (this.running && this.substeps > 0) {
      sim.run(solid, this.substeps);
      this.steps += this.substeps;
    }
    if (this.ws.buffered > 2 * this.msg.length) return;
    sim.packVolume(this.vol, null);
    const b = this.msg;
    b.writeUInt32LE(FRAME_MAGIC, 0);
    b.writeUInt32LE(sim.nx, 4);
    b.writeUInt32LE(sim.ny, 8);
    b.writeUInt32LE(sim.nz, 12);
    b.writeUInt32LE(this.steps >>> 0, 16);
    b.writeUInt32LE(this.frames++ >>> 0, 20);
    [solid.cx, solid.cy, solid.cz, solid.r, solid.vx, solid.vy, solid.vz]
      .forEach((v, j) => b.writeDoubleLE(v, 24 + 8 * j));
    this.ws.send(b);
  }

  close() {
    if (this.closed) return;
    this.closed = true;
    if (this.timer) clearImmediate(this.timer);
    if (this.scene) this.scene.sim.dispose();
    this.scene = null;
  }
}

// Starts a frame server; resolves to { server, port, close() }.
export function createFrameServer({ port = 8765, host = '127.0.0.1' } = {}) {
  const sessions = new Set();
  const server = http.createServer((req, res) => {
    res.writeHead(426, { 'Content-Type': 'text/plain' });
    res.end('WebSocket frame server: connect with remote_worker.js\n');
  });
  attachServer(server, (ws) => {
    const s = new Session(ws);
    sessions.add(s);
    ws.on('close', () => sessions.delete(s));
  });
  return new Promise((resolve) => {
    server.listen(port, host, () => resolve({
      server,
      port: server.address().port,
      close() {
        for (const s of sessions) { s.ws.close(); s.close(); }
        return new Promise((r) => server.close(r));
      },
    }));
  });
}

if (process.argv[1] && import.meta.url === pathToFileURL(process.argv[1]).href) {
  const arg = (k, d) => { const j = process.argv.indexOf(k); return j > 0 ? process.argv[j + 1] : d; };
  createFrameServer({ port: Number(arg('--port', 8765)), host: arg('--host', '127.0.0.1') })
    .then(({ port }) => console.log(`frame server on ws://${arg('--host', '127.0.0.1')}:${port}`));
}
