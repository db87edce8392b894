// Test driver: the addon checks caller array lengths against the context's counts
// before it calls the library (built against tests/abi_stub, whose contexts report
// 1 sampled voltage, 2 sampled currents, 3 probes and 2 NTFF frequencies).
//
//   node lengths_driver.mjs <addon.node> <out.json>
//
// Each case records the error message (null when the call went through).

import fs from 'fs';
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const [addonPath, outPath] = process.argv.slice(2);
const A = require(addonPath);

function message(fn) {
  try { fn(); return null; } catch (e) { return `${e.name}: ${e.message}`; }
}

const h = A.init({ nx: 8, ny: 8, nz: 8, dx: 1e-3, dy: 1e-3, dz: 1e-3, dt: 1e-12, numberOfTimeSteps: 40,
                   precision: 0, boundary: 0 });
const F = (n) => new Float64Array(n);
const face = (nc, profLen) => ({ nc, m_start: 1, e_start: 1, b_e: F(profLen), a_e: F(nc), b_m: F(nc), a_m: F(nc),
                                 CPsi0: F(4), CPsi1: F(4), CPsi2: F(4), CPsi3: F(4) });
const rec = {
  runShortV: message(() => A.runBatch(h, 0, 10, F(9), F(20))),
  runShortI: message(() => A.runBatch(h, 0, 10, F(10), F(19))),
  runOk: message(() => A.runBatch(h, 0, 10, F(10), F(20))),
  runNull: message(() => A.runBatch(h, 0, 10, null, null)),
  tracesShort: message(() => A.readTraces(h, 0, 5, F(5), F(9))),
  tracesOk: message(() => A.readTraces(h, 0, 5, F(5), F(10))),
  probesShort: message(() => A.readProbes(h, 0, 4, F(11))),
  probesOk: message(() => A.readProbes(h, 0, 4, F(12))),
  cpmlShort: message(() => A.uploadCpml(h, 0, face(3, 2))),
  cpmlOk: message(() => A.uploadCpml(h, 0, face(3, 3))),
  dirFreq: message(() => A.directivity(h, F(3), null, F(4), F(4), F(12), F(12))),
  dirOk: message(() => A.directivity(h, F(2), null, F(4), F(4), F(8), F(8))),
};
A.destroy(h);
fs.writeFileSync(outPath, JSON.stringify(rec));
