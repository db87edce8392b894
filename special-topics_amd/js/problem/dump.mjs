// Binary problem file: the FP64 problem as plain little-endian arrays plus a JSON
// header, so that non-JavaScript consumers (the Python harness that drives the
// native library through its C ABI) see exactly the arrays this builder produced.
//
// Layout: "FDTDPRB1" | u32 header bytes | u32 0 | header JSON (space padded to a
// multiple of 64) | arrays, each 64-byte aligned, offsets relative to file start.

import fs from 'fs';
import { COEFF_ORDER } from './grid.mjs';
import { FACE_TERMS, FACE_NAMES } from './cpml.mjs';
import { currentLoopTerms } from './samplers.mjs';
import { FF_ORDER, phaseTable } from './farfield.mjs';

const DTYPE = { Float64Array: 'f64', Float32Array: 'f32', Int32Array: 'i32', Uint8Array: 'u8' };

function bytesOf(arr) {
  return Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength);
}

/**
 * Collect every array the time loop consumes.  opts.fp32 stores coefficient and CPsi
 * arrays as Float32 (the fast path casts them anyway); everything else stays FP64.
 */
export function problemArrays(problem, opts) {
  const o = opts || {};
  const cast = (a) => (o.fp32 ? Float32Array.from(a) : a);
  const { grid, coeffs, cpml, samplers, sources } = problem;
  const ff = problem.farfield || problem.ff;
  const arrays = {};
  const header = {
    format: 'fdtd-problem/1',
    grid: {
      nx: grid.nx, ny: grid.ny, nz: grid.nz, dx: grid.dx, dy: grid.dy, dz: grid.dz, dt: grid.dt,
      min_x: grid.min_x, min_y: grid.min_y, min_z: grid.min_z,
      numberOfTimeSteps: grid.numberOfTimeSteps,
    },
    coeff_dtype: o.fp32 ? 'f32' : 'f64',
    cpml: [], vsources: [], svoltages: [], scurrents: [], probes: [], ff: { nFreq: 0 },
    ports: problem.ports || [], dft: problem.dft || null, meta: problem.meta || {},
  };
  if (problem.materialGrid) arrays.mat = problem.materialGrid;
  for (const name of COEFF_ORDER) arrays['coeff/' + name] = cast(coeffs[name]);

  for (const face of FACE_NAMES) {
    const f = cpml[face];
    if (!f) continue;
    const axis = face[0];
    const key = axis === 'x' ? 'i' : axis === 'y' ? 'j' : 'k';
    header.cpml.push({ face, nc: f.nc, m_start: f[key + '_m_start'], e_start: f[key + '_e_start'],
                       ascending: f.ascending ? 1 : 0 });
    for (const p of ['b_e', 'a_e', 'b_m', 'a_m']) arrays[`cpml/${face}/${p}`] = f[p];
    const t = FACE_TERMS[axis];
    [t.h[0], t.h[1], t.e[0], t.e[1]].forEach((term, slot) => {
      arrays[`cpml/${face}/CPsi${slot}`] = cast(f['CPsi_' + term[3]]);
    });
  }

  (sources.voltageSources || []).forEach((vs, n) => {
    const dir = 'xyz'.indexOf(vs.direction[0]);
    const coef = dir === 0 ? vs.Cexs : dir === 1 ? vs.Ceys : vs.Cezs;
    header.vsources.push({ dir, count: vs.field_indices.length });
    arrays[`vsrc/${n}/idx`] = vs.field_indices;
    arrays[`vsrc/${n}/coef`] = coef;
    arrays[`vsrc/${n}/vpe`] = vs.voltage_per_e_field;
  });
  (samplers.sampledVoltages || []).forEach((sv, n) => {
    header.svoltages.push({ dir: 'xyz'.indexOf(sv.direction[0]), count: sv.field_indices.length,
                            Csvf: sv.Csvf, label: sv.label || null });
    arrays[`sv/${n}/idx`] = sv.field_indices;
  });
  (samplers.sampledCurrents || []).forEach((si, n) => {
    const t = currentLoopTerms(si, grid);
    header.scurrents.push({ count: t.comp.length, negate: t.negate, label: si.label || null });
    arrays[`si/${n}/comp`] = t.comp;
    arrays[`si/${n}/index`] = t.index;
    arrays[`si/${n}/weight`] = t.weight;
  });
  // Point field probes, E then H, each as the 0-based cell (is-1, js-1, ks-1) and its
  // component (0..2 = x, y, z; 3 = magnitude 'm').
  [[samplers.sampledEFields || [], 0], [samplers.sampledHFields || [], 1]].forEach(([list, magnetic]) => {
    for (const pr of list) {
      const component = 'xyzm'.indexOf(pr.component);
      if (component < 0) throw new Error(`field probe component "${pr.component}" is not x, y, z or m`);
      header.probes.push({ magnetic, component, i: pr.is - 1, j: pr.js - 1, k: pr.ks - 1, label: pr.label || null });
    }
  });
  if (ff && ff.nFreq > 0) {
    header.ff = {
      nFreq: ff.nFreq, li: ff.li, lj: ff.lj, lk: ff.lk, ui: ff.ui, uj: ff.uj, uk: ff.uk,
      ni_ff: ff.ni_ff, nj_ff: ff.nj_ff, nk_ff: ff.nk_ff, frequencies: Array.from(ff.frequencies),
      order: FF_ORDER,
    };
    arrays['ff/w'] = ff.w;
    if (!o.noPhase) arrays['ff/phase'] = phaseTable(ff, grid.dt, grid.numberOfTimeSteps);
  }
  return { header, arrays };
}

export function writeProblemFile(problem, file, opts) {
  const { header, arrays } = problemArrays(problem, opts);
  const names = Object.keys(arrays);
  header.arrays = {};
  // Two passes: offsets depend on the header length, which depends on the offsets.
  let headerLen = 0;
  for (let pass = 0; ; pass++) {
    let off = 16 + headerLen;
    for (const name of names) {
      const a = arrays[name];
      header.arrays[name] = [DTYPE[a.constructor.name], off, a.length];
      off = Math.ceil((off + a.byteLength) / 64) * 64;
    }
    const need = Math.ceil((16 + Buffer.byteLength(JSON.stringify(header))) / 64) * 64 - 16;
    if (need <= headerLen) break;
    if (pass > 8) throw new Error('problem header layout did not converge');
    headerLen = need;
  }
  let text = JSON.stringify(header);
  text = text + ' '.repeat(headerLen - Buffer.byteLength(text));
  const fd = fs.openSync(file, 'w');
  try {
    const pre = Buffer.alloc(16);
    pre.write('FDTDPRB1', 0, 'latin1');
    pre.writeUInt32LE(headerLen, 8);
    fs.writeSync(fd, pre, 0, 16, 0);
    fs.writeSync(fd, Buffer.from(text, 'utf8'), 0, headerLen, 16);
    for (const name of names) {
      const [, off] = header.arrays[name];
      const b = bytesOf(arrays[name]);
      let done = 0;
      while (done < b.length) {
        const chunk = Math.min(b.length - done, 1 << 30);
        done += fs.writeSync(fd, b, done, chunk, off + done);
      }
    }
  } finally {
    fs.closeSync(fd);
  }
  return header;
}
