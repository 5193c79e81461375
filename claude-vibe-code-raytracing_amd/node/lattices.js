// Lattice tables with the exports of the reference's src/lattices.js
// (lattices.js:16-104), filled from the native library so JavaScript and the
// kernels share one table.
import { native } from './native.js';

const T = native.latticeTables();

export const Q27 = 27;
export const Q7 = 7;
export const EX27 = Int8Array.from(T.EX27);
export const EY27 = Int8Array.from(T.EY27);
export const EZ27 = Int8Array.from(T.EZ27);
export const W27 = Float32Array.from(T.W27);
export const OPP27 = Int8Array.from(T.OPP27);
export const EX7 = Int8Array.from(T.EX7);
export const EY7 = Int8Array.from(T.EY7);
export const EZ7 = Int8Array.from(T.EZ7);
export const W7 = Float32Array.from(T.W7);
export const OPP7 = Int8Array.from(T.OPP7);

export const CS2_27 = 1 / 3;
export const INV_CS2_27 = 3.0;
export const INV_2CS4_27 = 4.5;
export const CS2_7 = 1 / 4;
export const INV_CS2_7 = 4.0;

export const TAG_FLUID = 0;
export const TAG_SOLID = 1;
export const TAG_WALL = 2;
