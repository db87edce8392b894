// hip backend: the js_fdtd time loop on an AMD Instinct MI355X (gfx950).
//
// Implements the backend contract of js_fdtd/src/backends/types.js:27-32 and
// CONTRACT.md:104-139 -- { name, isAvailable(), async *run(problem), meta() } -- over
// the N-API addon fdtd_hip.node (csrc/napi_addon.c), which forwards to the C ABI of
// include/fdtd_hip.h.  Everything the reference runs per step (updateH, magnetic CPML,
// current capture, updateE with its i = nx ghost plane, electric CPML, source
// injection, voltage capture, far-field DFT accumulation; fdtdSolver.js:97-121) runs
// on the GPU; the host uploads the problem once, runs batches of `batchSize` steps,
// copies the per-step port traces into the samplers and, at the end, the far-field
// accumulators into `ff`.
//
// Precision: FP64 by default, bit-identical to the reference engine (no FMA, the
// reference's expression order and ghost-plane semantics).  options.precision = 'fp32'
// selects the fast path (FP32 fields and coefficients, FP64 traces and DFTs).
//
// isAvailable() never throws (registry.js:58-60): a missing or unloadable addon, or no
// device, reads as "not available".  run() throws when the addon is missing -- there
// is no CPU fallback inside this backend.

import { createRequire } from 'module';
import path from 'path';
import { fileURLToPath } from 'url';

import { problemArrays } from '../problem/dump.mjs';
import { FACE_NAMES } from '../problem/cpml.mjs';
import { FF_ORDER } from '../problem/farfield.mjs';
import { COEFF_ORDER, FIELD_ORDER } from '../problem/grid.mjs';

const require = createRequire(import.meta.url);
const HERE = path.dirname(fileURLToPath(import.meta.url));
export const ADDON_PATH = process.env.FDTD_HIP_ADDON || path.join(HERE, '..', 'fdtd_hip.node');

let addon = null;
let addonError = null;
let addonTried = false;

/** The loaded addon, or null (the reason is kept in addonLoadError()). */
export function loadAddon() {
  if (addonTried) return addon;
  addonTried = true;
  try {
    addon = require(ADDON_PATH);
  } catch (err) {
    addon = null;
    addonError = err;
  }
  return addon;
}

export function addonLoadError() {
  return addonError;
}

const PRECISION = { fp64: 0, f64: 0, fp32: 1, f32: 1 };
const BOUNDARY = { reference: 0, matlab: 1 };

function pick(v, fallback) {
  return v === undefined || v === null ? fallback : v;
}

function asInt32(a) {
  return a instanceof Int32Array ? a : Int32Array.from(a);
}

function asFloat64(a) {
  return a instanceof Float64Array ? a : Float64Array.from(a);
}

// FP32 under the reference boundary: the i = nx ghost-plane update makes long IFA runs
// grow without bound after about step 5000 (js_fdtd/scenarios/ifa-matlab-7000.json:3),
// and the growth amplifies FP32 rounding along with everything else, so such a run's S11
// can differ from the FP64 run by tens of dB.  Said once per process.
const FP32_REFERENCE_WARNING = "hip backend: precision 'fp32' with boundary 'reference': the reference's " +
  "i = nx ghost-plane update makes long runs diverge and amplifies FP32 rounding; use boundary 'matlab' " +
  "for physical FP32 results, or precision 'fp64' for the reference's own numbers";
let warnedFp32Reference = false;

/** Backend options: problem.options first, then FDTD_HIP_* environment variables. */
export function resolveOptions(options) {
  const o = options || {};
  const env = process.env;
  const precision = String(pick(o.precision, env.FDTD_HIP_PRECISION || 'fp64')).toLowerCase();
  const boundary = String(pick(o.boundary, env.FDTD_HIP_BOUNDARY || 'reference')).toLowerCase();
  if (!(precision in PRECISION)) throw new Error(`hip backend: unknown precision "${precision}" (fp64 | fp32)`);
  if (!(boundary in BOUNDARY)) throw new Error(`hip backend: unknown boundary "${boundary}" (reference | matlab)`);
  if (PRECISION[precision] === PRECISION.fp32 && boundary === 'reference' && !warnedFp32Reference) {
    warnedFp32Reference = true;
    console.warn(FP32_REFERENCE_WARNING);
  }
  let tuning = pick(o.tuning, env.FDTD_HIP_TUNING || 'auto');
  if (typeof tuning === 'string' && tuning !== 'auto' && tuning !== 'default') {
    const [p, c, s, z] = tuning.split(',').map(Number);
    tuning = { planesPerWorkgroup: p || 0, cellsPerThread: c || 0, coefficientStorage: s || 0, fuseCpml: z || 0 };
  }
  const batchSize = pick(o.batchSize, 50);
  if (!(Number.isInteger(batchSize) && batchSize > 0)) throw new Error(`hip backend: bad batchSize ${batchSize}`);
  return {
    batchSize, precision, boundary, tuning,
    portDftFrequencies: o.portDftFrequencies ? asFloat64(o.portDftFrequencies) : null,
    readFields: !!o.readFields,
    farfieldCuts: o.farfieldCuts || null,
    autotuneSteps: pick(o.autotuneSteps, 3),
  };
}

/**
 * Upload a js_fdtd problem (from js_fdtd/src/scenario.js:buildProblemFromScenario or
 * this package's builder) into a new device context; returns the handle and the
 * per-run bookkeeping.  The caller destroys the handle.
 */
export function uploadProblem(A, problem, opts) {
  const { grid } = problem;
  const ff = problem.ff || problem.farfield;
  const { header, arrays } = problemArrays(Object.assign({}, problem, { farfield: ff }), { noPhase: false });
  const h = A.init({
    nx: grid.nx, ny: grid.ny, nz: grid.nz, dx: grid.dx, dy: grid.dy, dz: grid.dz, dt: grid.dt,
    numberOfTimeSteps: grid.numberOfTimeSteps,
    precision: PRECISION[opts.precision], boundary: BOUNDARY[opts.boundary],
  });
  try {
    COEFF_ORDER.forEach((name, w) => A.uploadCoeff(h, w, asFloat64(arrays['coeff/' + name])));
    for (const f of header.cpml) {
      const p = (k) => asFloat64(arrays[`cpml/${f.face}/${k}`]);
      A.uploadCpml(h, FACE_NAMES.indexOf(f.face), {
        nc: f.nc, m_start: f.m_start, e_start: f.e_start,
        b_e: p('b_e'), a_e: p('a_e'), b_m: p('b_m'), a_m: p('a_m'),
        CPsi0: p('CPsi0'), CPsi1: p('CPsi1'), CPsi2: p('CPsi2'), CPsi3: p('CPsi3'),
      });
    }
    header.vsources.forEach((vs, n) => {
      A.uploadVSource(h, vs.dir, asInt32(arrays[`vsrc/${n}/idx`]), asFloat64(arrays[`vsrc/${n}/coef`]),
                      asFloat64(arrays[`vsrc/${n}/vpe`]));
    });
    header.svoltages.forEach((sv, n) => A.uploadSVoltage(h, sv.dir, asInt32(arrays[`sv/${n}/idx`]), sv.Csvf));
    header.scurrents.forEach((si, n) => {
      A.uploadSCurrent(h, asInt32(arrays[`si/${n}/comp`]), asInt32(arrays[`si/${n}/index`]),
                       asFloat64(arrays[`si/${n}/weight`]), si.negate ? 1 : 0);
    });
    for (const pr of header.probes) A.uploadFieldProbe(h, pr.magnetic, pr.component, pr.i, pr.j, pr.k);
    if (header.ff.nFreq > 0) {
      const b = header.ff;
      A.uploadFarfield(h, b.nFreq, Int32Array.of(b.li, b.lj, b.lk, b.ui, b.uj, b.uk), arrays['ff/phase']);
    }
    if (opts.portDftFrequencies && opts.portDftFrequencies.length > 0) {
      A.uploadPortDft(h, opts.portDftFrequencies);
    }
  } catch (err) {
    A.destroy(h);
    throw err;
  }
  return { h, header, ff };
}

function applyTuning(A, h, tuning, steps) {
  if (tuning === 'auto') return A.autotune(h, steps);
  if (tuning && typeof tuning === 'object') A.setTuning(h, tuning);
  return A.getTuning(h);
}

function snapshot(ts, total, startTime, sampledVoltages, dt) {
  return {
    step: ts + 1,
    total,
    elapsed: (Date.now() - startTime) / 1000,
    percent: (((ts + 1) / total) * 100).toFixed(1),
    voltage: sampledVoltages.length > 0 ? sampledVoltages[0].sampled_value[ts] : null,
    time_ns: (ts + 0.5) * dt * 1e9,
    backend: 'hip',
  };
}

/**
 * Radiated power and directivity cuts from the accumulators still on the device
 * (calcRadiatedPower / calcDirectivity of farfield.js, evaluated by the GPU; equal to
 * the sequential FP64 JavaScript to rounding).  cuts: [{theta, phi}].
 */
function farfieldOnDevice(A, h, ffHeader, cuts) {
  const nf = ffHeader.nFreq;
  const freqs = Float64Array.from(ffHeader.frequencies);
  const radPower = new Float64Array(nf);
  A.radiatedPower(h, radPower);
  return {
    radPower,
    cuts: cuts.map((cut) => {
      const th = asFloat64(cut.theta), ph = asFloat64(cut.phi), na = th.length;
      const t = new Float64Array(nf * na), p = new Float64Array(nf * na);
      A.directivity(h, freqs, radPower, th, ph, t, p);
      const rows = (x) => Array.from({ length: nf }, (_, f) => x.subarray(f * na, (f + 1) * na));
      return { dataTheta: rows(t), dataPhi: rows(p) };
    }),
  };
}

/**
 * After the steps have run: the NTFF accumulators (and, with farfieldCuts, the GPU
 * far-field post), the running port DFTs and, with readFields, the fields.
 */
function collectResults(A, h, problem, opts, header, ff) {
  const { grid, samplers } = problem;
  const sampledVoltages = (samplers && samplers.sampledVoltages) || [];
  const sampledCurrents = (samplers && samplers.sampledCurrents) || [];
  const nsv = sampledVoltages.length, nsi = sampledCurrents.length;
  if (header.ff.nFreq > 0) {
    FF_ORDER.forEach((name, q) => A.readFarfield(h, q, ff[name].re, ff[name].im));
    if (opts.farfieldCuts) problem.farfieldPost = farfieldOnDevice(A, h, header.ff, opts.farfieldCuts);
  }
  if (opts.portDftFrequencies && opts.portDftFrequencies.length > 0) {
    const m = opts.portDftFrequencies.length;
    const re = new Float64Array((nsv + nsi) * m), im = new Float64Array((nsv + nsi) * m);
    A.readPortDft(h, re, im);
    const part = (o) => ({ re: re.slice(o * m, (o + 1) * m), im: im.slice(o * m, (o + 1) * m) });
    problem.portDft = {
      frequencies: opts.portDftFrequencies,
      voltages: sampledVoltages.map((_, o) => part(o)),
      currents: sampledCurrents.map((_, o) => part(nsv + o)),
    };
  }
  if (opts.readFields && grid.fields) {
    FIELD_ORDER.forEach((name, q) => { if (grid.fields[name]) A.readField(h, q, grid.fields[name]); });
  }
}

// The point field probes in upload order (E, then H), as dump.mjs lists them.
function probeList(problem) {
  const { samplers } = problem;
  return [...((samplers && samplers.sampledEFields) || []), ...((samplers && samplers.sampledHFields) || [])];
}

// Copy probe traces of steps [start, start+count) into each probe's sampled_value.
function readProbes(A, h, probes, start, count, buf) {
  if (probes.length === 0) return;
  const p = buf.subarray(0, probes.length * count);
  A.readProbes(h, start, count, p);
  probes.forEach((o, n) => o.sampled_value.set(p.subarray(n * count, (n + 1) * count), start));
}

function requireAddon() {
  const A = loadAddon();
  if (!A) {
    const why = addonError ? addonError.message : 'not found';
    throw new Error(`hip addon ${ADDON_PATH} could not be loaded (${why}); build it with __graft_entry__.build()`);
  }
  return A;
}

let lastRun = null;

/** @type {import('./types.js').Backend} */
const backend = {
  name: 'hip',

  async isAvailable() {
    try {
      const A = loadAddon();
      if (!A || typeof A.deviceCount !== 'function') return false;
      return A.deviceCount() > 0;
    } catch (err) {
      return false;
    }
  },

  meta() {
    return Object.assign({ name: 'hip', device: 'gfx950', precision: 'f64 (default) | f32',
                           farfield: 'gpu' }, lastRun || {});
  },

  async *run(problem) {
    const A = requireAddon();
    const opts = resolveOptions(problem.options);
    const { grid, samplers } = problem;
    const total = grid.numberOfTimeSteps;
    const sampledVoltages = (samplers && samplers.sampledVoltages) || [];
    const sampledCurrents = (samplers && samplers.sampledCurrents) || [];
    const probes = probeList(problem);
    const tSetup = Date.now();
    const { h, header, ff } = uploadProblem(A, problem, opts);
    try {
      const tuning = applyTuning(A, h, opts.tuning, opts.autotuneSteps);
      lastRun = { precision: opts.precision, boundary: opts.boundary, tuning,
                  setup_s: (Date.now() - tSetup) / 1000 };
      const nsv = sampledVoltages.length, nsi = sampledCurrents.length;
      const vBuf = new Float64Array(Math.max(nsv, 1) * opts.batchSize);
      const iBuf = new Float64Array(Math.max(nsi, 1) * opts.batchSize);
      const pBuf = new Float64Array(Math.max(probes.length, 1) * opts.batchSize);
      const startTime = Date.now();
      for (let start = 0; start < total; start += opts.batchSize) {
        const count = Math.min(opts.batchSize, total - start);
        const v = vBuf.subarray(0, Math.max(nsv, 1) * count);
        const c = iBuf.subarray(0, Math.max(nsi, 1) * count);
        A.runBatch(h, start, count, nsv ? v : null, nsi ? c : null);
        // Traces come back observer-major: [observer][step of the batch].
        for (let o = 0; o < nsv; o++) sampledVoltages[o].sampled_value.set(v.subarray(o * count, (o + 1) * count), start);
        for (let o = 0; o < nsi; o++) sampledCurrents[o].sampled_value.set(c.subarray(o * count, (o + 1) * count), start);
        readProbes(A, h, probes, start, count, pBuf);
        yield snapshot(start + count - 1, total, startTime, sampledVoltages, grid.dt);
        await new Promise((r) => setTimeout(r, 0));
      }
      collectResults(A, h, problem, opts, header, ff);
      lastRun.run_s = (Date.now() - startTime) / 1000;
    } finally {
      A.destroy(h);
    }
  },

  /**
   * Many problems (of any grid size) run together: one launch per phase covers all of
   * them (fdtd_hip_multi_*).  Every problem gets the same results as from run(); the
   * precision, step count and tuning come from the first problem's options, which
   * must agree on precision and step count.  Resolves once all have finished.
   */
  async runMany(problems) {
    const A = requireAddon();
    if (!Array.isArray(problems) || problems.length === 0) throw new Error('hip backend: runMany needs problems');
    const opts = problems.map((p) => resolveOptions(p.options));
    const total = problems[0].grid.numberOfTimeSteps;
    problems.forEach((p, n) => {
      if (opts[n].precision !== opts[0].precision) throw new Error('hip backend: runMany needs one precision');
      if (p.grid.numberOfTimeSteps !== total) throw new Error('hip backend: runMany needs one step count');
    });
    const tSetup = Date.now();
    const uploaded = [];
    let m = null;
    try {
      for (let n = 0; n < problems.length; n++) uploaded.push(uploadProblem(A, problems[n], opts[n]));
      const t = opts[0].tuning;
      m = A.multiCreate(uploaded.map((u) => u.h), t && typeof t === 'object' ? t : null);
      const tuning = t === 'auto' ? A.multiAutotune(m, opts[0].autotuneSteps) : null;
      lastRun = { precision: opts[0].precision, problems: problems.length, tuning,
                  setup_s: (Date.now() - tSetup) / 1000 };
      const startTime = Date.now();
      A.multiRun(m, 0, total);
      A.multiSynchronize(m);
      problems.forEach((p, n) => {
        const { h, header, ff } = uploaded[n];
        const sv = (p.samplers && p.samplers.sampledVoltages) || [];
        const si = (p.samplers && p.samplers.sampledCurrents) || [];
        const v = new Float64Array(Math.max(sv.length, 1) * total);
        const c = new Float64Array(Math.max(si.length, 1) * total);
        A.readTraces(h, 0, total, sv.length ? v : null, si.length ? c : null);
        sv.forEach((o, k) => o.sampled_value.set(v.subarray(k * total, (k + 1) * total)));
        si.forEach((o, k) => o.sampled_value.set(c.subarray(k * total, (k + 1) * total)));
        const probes = probeList(p);
        readProbes(A, h, probes, 0, total, new Float64Array(Math.max(probes.length, 1) * total));
        collectResults(A, h, p, opts[n], header, ff);
      });
      lastRun.run_s = (Date.now() - startTime) / 1000;
    } finally {
      if (m) A.multiDestroy(m);
      for (const u of uploaded) A.destroy(u.h);
    }
  },
};

export default backend;
