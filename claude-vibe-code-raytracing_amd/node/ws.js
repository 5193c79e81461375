// Minimal WebSocket (RFC 6455) for the frame transport: server upgrade and a
// client, text and binary messages, fragmentation, ping/pong and close.  No
// extensions.  Node 12, no dependencies.
import crypto from 'crypto';
import http from 'http';
import { EventEmitter } from 'events';

const GUID = '258EAFA5-E914-47DA-95CA-C5AB0DC85B11';

export function acceptKey(key) {
  return crypto.createHash('sha1').update(key + GUID).digest('base64');
}

function header(opcode, len, masked) {
  let h;
  if (len < 126) {
    h = Buffer.alloc(2);
    h[1] = len;
  } else if (len < 65536) {
    h = Buffer.alloc(4);
    h[1] = 126;
    h.writeUInt16BE(len, 2);
  } else {
    h = Buffer.alloc(10);
    h[1] = 127;
    h.writeUInt32BE(Math.floor(len / 4294967296), 2);
    h.writeUInt32BE(len >>> 0, 6);
  }
  h[0] = 0x80 | opcode;
  if (masked) h[1] |= 0x80;
  return h;
}

// One frame as a list of buffers (header, [mask key], payload).  Client frames
// are masked (RFC 6455 5.3).
export function encodeFrame(opcode, payload, masked) {
  const h = header(opcode, payload.length, masked);
  if (!masked) return [h, payload];
  const key = crypto.randomBytes(4);
  const p = Buffer.allocUnsafe(payload.length);
  for (let i = 0; i < payload.length; i++) p[i] = payload[i] ^ key[i & 3];
  return [h, key, p];
}

// Incremental frame parser: push() bytes, get back complete messages.
class Parser {
  constructor() {
    this.buf = Buffer.alloc(0);
    this.parts = [];
    this.opcode = 0;
  }

  push(chunk) {
    this.buf = this.buf.length ? Buffer.concat([this.buf, chunk]) : chunk;
    const out = [];
    for (;;) {
      const b = this.buf;
      if (b.length < 2) break;
      const fin = (b[0] & 0x80) !== 0;
      const op = b[0] & 0x0f;
      const masked = (b[1] & 0x80) !== 0;
      let len = b[1] & 0x7f;
      let off = 2;
      if (len === 126) {
        if (b.length < 4) break;
        len = b.readUInt16BE(2);
        off = 4;
      } else if (len === 127) {
        if (b.length < 10) break;
        len = b.readUInt32BE(2) * 4294967296 + b.readUInt32BE(6);
        off = 10;
      }
      const keyOff = off;
      if (masked) off += 4;
      if (b.length < off + len) break;
      let payload = b.slice(off, off + len);
      if (masked) {
        const p = Buffer.allocUnsafe(len);
        for (let i = 0; i < len; i++) p[i] = payload[i] ^ b[keyOff + (i & 3)];
        payload = p;
      }
      this.buf = b.slice(off + len);
      if (op >= 8) {
        out.push({ opcode: op, payload });  // control frames are never fragmented
        continue;
      }
      if (op !== 0) this.opcode = op;
      this.parts.push(payload);
      if (fin) {
        out.push({ opcode: this.opcode, payload: this.parts.length === 1 ? this.parts[0] : Buffer.concat(this.parts) });
        this.parts = [];
      }
    }
    return out;
  }
}

// An open WebSocket over a net.Socket.  Events: 'message' (data, isBinary),
// 'close'.
export class WsConnection extends EventEmitter {
  constructor(socket, isClient, head) {
    super();
    this.socket = socket;
    this.isClient = isClient;
    this.parser = new Parser();
    this.closed = false;
    socket.setNoDelay(true);
    socket.on('data', (d) => this._data(d));
    socket.on('close', () => this._closed());
    socket.on('error', () => this._closed());
    if (head && head.length) this._data(head);
  }

  _data(d) {
    for (const m of this.parser.push(d)) {
      if (m.opcode === 1) this.emit('message', m.payload.toString('utf8'), false);
      else if (m.opcode === 2) this.emit('message', m.payload, true);
      else if (m.opcode === 8) this.close();
      else if (m.opcode === 9) this._write(10, m.payload);
    }
  }

  _write(opcode, payload) {
    if (this.closed) return false;
    let ok = true;
    for (const part of encodeFrame(opcode, payload, this.isClient)) ok = this.socket.write(part);
    return ok;
  }

  // Strings go as text, Buffers / typed arrays as binary.  Returns false when
  // the socket's buffer is full (the caller may skip frames).
  send(data) {
    if (typeof data === 'string') return this._write(1, Buffer.from(data, 'utf8'));
    const b = Buffer.isBuffer(data) ? data : Buffer.from(data.buffer, data.byteOffset, data.byteLength);
    return this._write(2, b);
  }

  // Bytes queued in the socket and not yet sent.
  get buffered() { return this.socket.writableLength; }

  close() {
    if (this.closed) return;
    this._write(8, Buffer.alloc(0));
    this.closed = true;
    this.socket.end();
    this.emit('close');
  }

  _closed() {
    if (this.closed) return;
    this.closed = true;
    this.emit('close');
  }
}

// Accepts WebSocket upgrades on an http.Server; onConnection(ws, request).
export function attachServer(server, onConnection) {
  server.on('upgrade', (req, socket, head) => {
    const key = req.headers['sec-websocket-key'];
    if (!key || String(req.headers.upgrade).toLowerCase() !== 'websocket') {
      socket.end('HTTP/1.1 400 Bad Request\r\n\r\n');
      return;
    }
    socket.write('HTTP/1.1 101 Switching Protocols\r\nUpgrade: websocket\r\nConnection: Upgrade\r\n' +
                 `Sec-WebSocket-Accept: ${acceptKey(key)}\r\n\r\n`);
    onConnection(new WsConnection(socket, false, head), req);
  });
}

// Client: resolves to an open WsConnection.
export function connect(url) {
  return new Promise((resolve, reject) => {
    const u = new URL(url);
    const key = crypto.randomBytes(16).toString('base64');
    const req = http.request({
      host: u.hostname, port: u.port, path: u.pathname + u.search,
      headers: { Connection: 'Upgrade', Upgrade: 'websocket', 'Sec-WebSocket-Version': '13',
                 'Sec-WebSocket-Key': key },
    });
    req.on('upgrade', (res, socket, head) => {
      if (res.headers['sec-websocket-accept'] !== acceptKey(key)) {
        socket.destroy();
        reject(new Error('bad Sec-WebSocket-Accept'));
        return;
      }
      resolve(new WsConnection(socket, true, head));
    });
    req.on('response', (res) => reject(new Error(`upgrade refused: ${res.statusCode}`)));
    req.on('error', reject);
    req.end();
  });
}
