// ws.js on the CPU: an echo server and a client exchange text and binary
// messages of every length class (7-bit, 16-bit and 64-bit lengths), masked
// client frames, and fragmented frames written by hand.
import http from 'http';
import { attachServer, connect, encodeFrame } from '../../claude-vibe-code-raytracing_amd/node/ws.js';

function fail(m) { console.error('FAIL', m); process.exit(1); }

async function main() {
  const server = http.createServer();
  attachServer(server, (ws) => ws.on('message', (d, bin) => ws.send(bin ? d : `echo:${d}`)));
  await new Promise((r) => server.listen(0, '127.0.0.1', r));
  const port = server.address().port;
  const c = await connect(`ws://127.0.0.1:${port}/`);
  const got = [];
  let wake = null;
  c.on('message', (d, bin) => { got.push({ d, bin }); if (wake) wake(); });
  const next = () => new Promise((r) => { if (got.length) r(got.shift()); else wake = () => { wake = null; r(got.shift()); }; });
  c.send('hello');
  let m = await next();
  if (m.bin || m.d !== 'echo:hello') fail('text');
  for (const len of [0, 1, 125, 126, 1000, 65535, 65536, 300000]) {
    const b = Buffer.alloc(len);
    for (let i = 0; i < len; i++) b[i] = (i * 131 + len) & 255;
    c.send(new Uint8Array(b.buffer, b.byteOffset, b.length));
    m = await next();
    if (!m.bin || m.d.length !== len || !m.d.equals(b)) fail(`binary ${len}`);
  }
  // A fragmented text message: "frag" + "ment", written as two raw frames.
  const f1 = encodeFrame(1, Buffer.from('frag'), true);
  f1[0][0] &= 0x7f;  // FIN cleared
  const f2 = encodeFrame(0, Buffer.from('ment'), true);
  for (const part of f1.concat(f2)) c.socket.write(part);
  m = await next();
  if (m.d !== 'echo:fragment') fail('fragmented');
  c.close();
  server.close();
  console.log('WS OK');
}

main().catch((e) => fail(e.stack || String(e)));
