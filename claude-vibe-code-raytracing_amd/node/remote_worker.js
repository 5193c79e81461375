// Drop-in for the page's simulation worker (src/worker.js) when the simulation
// runs on a GPU behind frame_server.js.  The page keeps main.js and the
// Renderer unchanged: this module worker accepts the same messages and posts
// the same {type:'frame', vol, solid} messages, whose vol goes straight to
// Renderer.uploadVolumeRaw (src/render.js:430-436).  Deploy it as worker.js, or
// point main.js's Worker at it; the server URL comes from the worker URL's
// ?server= parameter (default ws://127.0.0.1:8765).
const FRAME_MAGIC = 0x46384752;
const FRAME_HEADER = 80;

const scope = typeof self !== 'undefined' ? self : globalThis;
const search = scope.location && scope.location.search ? scope.location.search : '';
const url = new URLSearchParams(search).get('server') || 'ws://127.0.0.1:8765';

let ws = null;
let open = false;
const outbox = [];
let pool = [];
let volBytes = 0;

function connect() {
  ws = new scope.WebSocket(url);
  ws.binaryType = 'arraybuffer';
  ws.onopen = () => {
    open = true;
    for (const m of outbox) ws.send(m);
    outbox.length = 0;
  };
  ws.onmessage = (ev) => {
    if (typeof ev.data === 'string') return;
    onFrame(ev.data);
  };
  ws.onclose = () => { open = false; };
}

function send(m) {
  const s = JSON.stringify(m);
  if (open) ws.send(s);
  else outbox.push(s);
}

function onFrame(data) {
  const dv = new DataView(data);
  if (dv.getUint32(0, true) !== FRAME_MAGIC) return;
  const n = dv.getUint32(4, true) * dv.getUint32(8, true) * dv.getUint32(12, true);
  if (2 * n !== volBytes) { volBytes = 2 * n; pool = []; }
  const vol = pool.length ? pool.pop() : new Uint8Array(volBytes);
  vol.set(new Uint8Array(data, FRAME_HEADER, volBytes));
  const f = (j) => dv.getFloat64(24 + 8 * j, true);
  scope.postMessage({
    type: 'frame',
    vol,
    solid: { cx: f(0), cy: f(1), cz: f(2), r: f(3), vx: f(4), vy: f(5), vz: f(6) },
    steps: dv.getUint32(16, true),
  }, [vol.buffer]);
}

scope.onmessage = (e) => {
  const m = e.data;
  if (m.type === 'releaseBuf') {
    // The page hands a frame buffer back for reuse (src/worker.js:158-163).
    if (m.buffer && m.buffer.byteLength === volBytes) pool.push(new Uint8Array(m.buffer));
    return;
  }
  send(m);
};

connect();
