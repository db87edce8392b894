#!/usr/bin/env node
// Batch run from a scenario to results/out.json, as js_fdtd/src/index.js:14-135 does:
//
//   node cli.mjs [--scenario=NAME|PATH] [--backend=auto|hip] [--out=FILE]
//                [--precision=fp64|fp32] [--boundary=reference|matlab] [--steps=N]
//                [--farfield-post=gpu|js]
//
// FDTD_SCENARIO / FDTD_BACKEND / FDTD_STEPS are honoured like the reference CLI.  The
// time loop runs on the selected backend; the port DFTs, S-parameters and NTFF
// directivity cuts are computed afterwards in FP64 with the reference's formulas, and
// the output has the same schema (meta, sparams, farfield, timeDomain).  With
// --farfield-post=gpu (the default on the hip backend) the radiated power and the cuts
// are evaluated on the GPU from the device-resident accumulators, equal to the
// JavaScript to rounding; --farfield-post=js reproduces the reference bit for bit.

import fs from 'fs';
import path from 'path';

import { loadScenario, listScenarios, buildProblemFromScenario } from './problem/scenario.mjs';
import { selectBackend, listBackends } from './backends/registry.mjs';
import { frequencyGrid, computeObserverDFT } from './post/dft.mjs';
import { postProcessSParameters } from './post/sparams.mjs';
import { calcRadiatedPower, calcDirectivity, standardCuts } from './post/farfield.mjs';

function arg(name, env, dflt) {
  const a = process.argv.find((x) => x.startsWith(`--${name}=`));
  if (a) return a.slice(name.length + 3);
  if (env && process.env[env]) return process.env[env];
  return dflt;
}

const toDb = (x) => 10 * Math.log10(Math.max(x, 1e-10));

/** Post-processing and the out.json record, as index.js:63-130 builds it. */
export function buildResults(sim, scenarioName, farfieldPost) {
  const { grid, samplers, farfield, ports, frequencies, numberOfTimeSteps, dt, dft } = sim;
  const freqArr = frequencyGrid(dft);
  computeObserverDFT(samplers.sampledVoltages, dt, 0, freqArr);
  computeObserverDFT(samplers.sampledCurrents, dt, -dt / 2, freqArr);
  postProcessSParameters({ ports, sampledVoltages: samplers.sampledVoltages,
                           sampledCurrents: samplers.sampledCurrents, dt, freqArr });
  const ffData = [];
  if (frequencies.length > 0) {
    const cuts = standardCuts();
    let xy, xz;
    if (farfieldPost) {
      [xy, xz] = farfieldPost.cuts;
    } else {
      const prad = calcRadiatedPower(farfield, grid);
      xy = calcDirectivity(farfield, grid, prad, cuts.xy.theta, cuts.xy.phi);
      xz = calcDirectivity(farfield, grid, prad, cuts.xz.theta, cuts.xz.phi);
    }
    const deg = Array.from(cuts.angles, (x) => x * 180 / Math.PI);
    for (let f = 0; f < frequencies.length; f++) {
      for (const [plane, d] of [['xy', xy], ['xz', xz]]) {
        ffData.push({ plane, freq_GHz: frequencies[f] / 1e9, angles_deg: deg.slice(),
                      dataTheta_dB: Array.from(d.dataTheta[f], toDb), dataPhi_dB: Array.from(d.dataPhi[f], toDb) });
      }
    }
  }
  const time_ns = [];
  for (let i = 0; i < numberOfTimeSteps; i++) time_ns.push((i + 0.5) * dt * 1e9);
  return {
    meta: {
      generated: new Date().toISOString(),
      scenario: (sim.meta && sim.meta.name) || scenarioName,
      description: (sim.meta && sim.meta.description) || null,
      numberOfTimeSteps, dt,
      frequencies_Hz: Array.from(freqArr),
    },
    sparams: ports.filter((p) => p.is_source_port).map((p) => ({
      label: 'S11',
      frequencies_GHz: Array.from(freqArr, (f) => f / 1e9),
      magnitude_dB: Array.from(p.S[0].mag_dB),
      phase_deg: Array.from(p.S[0].phase_deg),
    })),
    farfield: ffData,
    timeDomain: {
      voltages: samplers.sampledVoltages.map((v) => ({ label: v.label, time_ns: time_ns.slice(),
                                                       values: Array.from(v.sampled_value) })),
    },
  };
}

async function main() {
  const scenarioName = arg('scenario', 'FDTD_SCENARIO', 'ifa-dualband-baseline');
  const backendArg = arg('backend', 'FDTD_BACKEND', 'auto');
  const steps = Number(arg('steps', null, '0')) || undefined;
  const out = arg('out', null, path.resolve('results', 'out.json'));
  console.log(`Initializing FDTD problem from scenario "${scenarioName}"...`);
  console.log(`Available scenarios: ${listScenarios().join(', ')}`);
  const sim = buildProblemFromScenario(loadScenario(scenarioName), { steps });

  const probed = await listBackends();
  console.log('Available backends: ' + probed.map((b) => `${b.name}${b.available ? '' : '(x)'}`).join(', '));
  const { name, backend } = await selectBackend(backendArg);
  console.log(`Running FDTD on backend: ${name}`);

  const ffPost = arg('farfield-post', null, name === 'hip' ? 'gpu' : 'js');
  if (ffPost !== 'gpu' && ffPost !== 'js') throw new Error(`--farfield-post must be gpu or js, not ${ffPost}`);
  const cuts = standardCuts();
  const problem = {
    grid: sim.grid, coeffs: sim.coeffs, cpml: sim.cpml, samplers: sim.samplers, sources: sim.sources,
    ff: sim.farfield,
    options: { batchSize: 50, precision: arg('precision', 'FDTD_HIP_PRECISION', 'fp64'),
               boundary: arg('boundary', 'FDTD_HIP_BOUNDARY', 'reference'),
               farfieldCuts: ffPost === 'gpu' ? [cuts.xy, cuts.xz] : null },
  };
  const start = Date.now();
  for await (const snap of backend.run(problem)) {
    process.stdout.write(`\r[${snap.backend}] Step ${snap.step} / ${snap.total} - elapsed ${snap.elapsed.toFixed(1)}s    `);
  }
  const loopS = (Date.now() - start) / 1000;
  console.log(`\nSimulation complete in ${loopS}s. Processing records...`);
  const results = buildResults(sim, scenarioName, problem.farfieldPost);
  results.meta.backend = Object.assign({ loop_s: loopS }, backend.meta ? backend.meta() : { name });
  fs.mkdirSync(path.dirname(out), { recursive: true });
  fs.writeFileSync(out, JSON.stringify(results));
  console.log(`Results written to ${out}`);
}

const invokedDirectly = process.argv[1] && path.resolve(process.argv[1]) === new URL(import.meta.url).pathname;
if (invokedDirectly) {
  main().catch((err) => {
    console.error(err);
    process.exit(1);
  });
}
