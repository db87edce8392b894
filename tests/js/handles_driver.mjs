// Test driver: problem and batch handle lifetimes in the N-API addon.
//
//   node --expose-gc handles_driver.mjs <addon.node> <out.json> [live]
//
// 1. destroy() of a problem held by a live batch throws, and succeeds after
//    multiDestroy(); a second destroy() of the same handle throws.
// 2. A problem handle that becomes unreachable while a batch holds it is collected;
//    the batch still runs it, and multiDestroy() then releases it.
// With "live", each problem is the small golden scenario uploaded by the backend, and
// the batch really runs (needs the GPU); without it the driver only makes handles.

import fs from 'fs';
import path from 'path';
import { createRequire } from 'module';
import { fileURLToPath } from 'url';

const require = createRequire(import.meta.url);
const HERE = path.dirname(fileURLToPath(import.meta.url));
const [addonPath, outPath, mode] = process.argv.slice(2);
const A = require(addonPath);
const live = mode === 'live';

function message(fn) {
  try { fn(); return null; } catch (e) { return String(e.message); }
}

async function collect() {
  for (let r = 0; r < 6; r++) {
    global.gc();
    await new Promise((res) => setImmediate(res));
  }
}

async function main() {
  let uploadOne;
  if (live) {
    const { loadScenario, buildProblemFromScenario } = await import('../../special-topics_amd/js/problem/scenario.mjs');
    const { uploadProblem, resolveOptions } = await import('../../special-topics_amd/js/backends/hipBackend.mjs');
    const scen = path.join(HERE, '..', 'golden', 'scenarios', 'golden-small.json');
    uploadOne = () => {
      const sim = buildProblemFromScenario(loadScenario(scen), { steps: 40 });
      const problem = { grid: sim.grid, coeffs: sim.coeffs, cpml: sim.cpml, samplers: sim.samplers,
                        sources: sim.sources, ff: sim.farfield, options: {} };
      return uploadProblem(A, problem, resolveOptions({})).h;
    };
  } else {
    const dims = { nx: 8, ny: 8, nz: 8, dx: 1e-3, dy: 1e-3, dz: 1e-3, dt: 1e-12, numberOfTimeSteps: 40,
                   precision: 0, boundary: 0 };
    uploadOne = () => A.init(dims);
  }

  const rec = {};
  {
    const h1 = uploadOne(), h2 = uploadOne();
    const m = A.multiCreate([h1, h2], null);
    rec.refused = message(() => A.destroy(h1));
    A.multiRun(m, 0, 4);
    A.multiSynchronize(m);
    A.multiDestroy(m);
    rec.afterBatch = message(() => A.destroy(h1));
    rec.twice = message(() => A.destroy(h1));
    A.destroy(h2);
  }
  {
    // The first handle is made in a function frame that has returned, so only the
    // batch refers to its context.
    const keep = uploadOne();
    const make = () => A.multiCreate([uploadOne(), keep], null);
    const m = make();
    await collect();
    A.deviceCount();                      // marker: the handle has been collected
    rec.runAfterCollect = message(() => { A.multiRun(m, 0, 4); A.multiSynchronize(m); });
    A.multiDestroy(m);
    A.deviceCount();                      // marker: the batch is gone
    await collect();
    A.destroy(keep);
  }
  fs.writeFileSync(outPath, JSON.stringify(rec));
}

main().catch((err) => {
  console.error(err && err.stack ? err.stack : err);
  process.exit(1);
});
