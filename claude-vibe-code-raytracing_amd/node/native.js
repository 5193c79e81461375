// Loads the N-API addon (../lib/lbm.node, built by ../build.py).  There is no
// CPU fallback: without the addon and a GPU the import fails loudly.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
export const native = require('../lib/lbm.node');
