// Update coefficients for the six Yee components.
//
// General coefficients follow js_fdtd/src/coefficients.js:22-116 and the lumped
// element overrides coefficients.js:136-423 (voltage sources :140-189, resistors
// :243-283, capacitors :285-325).  The expression trees are the reference's, operand
// order included, because the FP64 drop-in must reproduce its coefficient arrays bit
// for bit; note in particular that the general denominator is 2*eps_r*eps0 + dt*sigma
// while the lumped one is 2*eps0*eps_r + dt*sigma + a_term.

import { EPS0, MU0 } from './constants.mjs';
import { COEFF_GROUPS, componentLength, boxIndexList } from './grid.mjs';

// Per component: the material arrays it averages and, for the two curl
// coefficients, [sign, cell-size axis] of the numerator 2*dt/d.
const GENERAL = {
  Ex: { eps: 'eps_r_x', sig: 'sigma_e_x', c1: [+1, 1], c2: [-1, 2] },
  Ey: { eps: 'eps_r_y', sig: 'sigma_e_y', c1: [+1, 2], c2: [-1, 0] },
  Ez: { eps: 'eps_r_z', sig: 'sigma_e_z', c1: [+1, 0], c2: [-1, 1] },
  Hx: { eps: 'mu_r_x', sig: 'sigma_m_x', c1: [+1, 2], c2: [-1, 1] },
  Hy: { eps: 'mu_r_y', sig: 'sigma_m_y', c1: [+1, 0], c2: [-1, 2] },
  Hz: { eps: 'mu_r_z', sig: 'sigma_m_z', c1: [+1, 1], c2: [-1, 0] },
};

export function computeGeneralCoefficients(mc, grid) {
  const { nx, ny, nz, dx, dy, dz, dt } = grid;
  const twoDt = [2 * dt / dx, 2 * dt / dy, 2 * dt / dz];
  const coeffs = {};
  for (const comp of Object.keys(GENERAL)) {
    const g = GENERAL[comp];
    const vac = comp[0] === 'E' ? EPS0 : MU0;
    const n = componentLength(comp, nx, ny, nz);
    const [nSelf, n1, n2] = COEFF_GROUPS[comp];
    const cs = new Float64Array(n), c1 = new Float64Array(n), c2 = new Float64Array(n);
    const num1 = g.c1[0] > 0 ? twoDt[g.c1[1]] : -twoDt[g.c1[1]];
    const num2 = g.c2[0] > 0 ? twoDt[g.c2[1]] : -twoDt[g.c2[1]];
    const rel = mc[g.eps], sig = mc[g.sig];
    for (let p = 0; p < n; p++) {
      const denom = 2 * rel[p] * vac + dt * sig[p];
      cs[p] = (2 * rel[p] * vac - dt * sig[p]) / denom;
      c1[p] = num1 / denom;
      c2[p] = num2 / denom;
    }
    coeffs[nSelf] = cs;
    coeffs[n1] = c1;
    coeffs[n2] = c2;
  }
  return coeffs;
}

// Lumped elements along axis a: the transverse axes in cyclic order, the E
// component's index list shape and its three coefficient arrays.
const LUMPED = [
  { comp: 'Ex', t1: 1, t2: 2, src: 'Cexs' },
  { comp: 'Ey', t1: 2, t2: 0, src: 'Ceys' },
  { comp: 'Ez', t1: 0, t2: 1, src: 'Cezs' },
];
const AXIS_OF = { x: 0, y: 1, z: 2 };

/** 1-based bounding-box indices of a lumped element (coefficients.js:142-151). */
export function elementIndices(el, grid) {
  const { dx, dy, dz, min_x, min_y, min_z } = grid;
  return {
    is: Math.round((el.min_x - min_x) / dx) + 1,
    js: Math.round((el.min_y - min_y) / dy) + 1,
    ks: Math.round((el.min_z - min_z) / dz) + 1,
    ie: Math.round((el.max_x - min_x) / dx) + 1,
    je: Math.round((el.max_y - min_y) / dy) + 1,
    ke: Math.round((el.max_z - min_z) / dz) + 1,
  };
}

/** Flat indices of the E edges an element along `axis` covers (k outer, i inner). */
export function elementFieldIndices(b, axis, grid) {
  const { ny, nz, nyp1, nzp1 } = grid;
  if (axis === 0) return boxIndexList(b.is, b.ie - 1, b.js, b.je, b.ks, b.ke, nyp1, nzp1);
  if (axis === 1) return boxIndexList(b.is, b.ie, b.js, b.je - 1, b.ks, b.ke, ny, nzp1);
  return boxIndexList(b.is, b.ie, b.js, b.je, b.ks, b.ke - 1, nyp1, nz);
}

/**
 * Overwrite the E coefficients on the edges of voltage sources, resistors and
 * capacitors.  kind: 'vsource' (attaches field_indices and C?s), 'resistor' or
 * 'capacitor' (attaches field_indices).
 */
function applyLumped(coeffs, mc, grid, el, kind) {
  const { dt } = grid;
  const d = [grid.dx, grid.dy, grid.dz];
  const axis = AXIS_OF[el.direction[0]];
  const L = LUMPED[axis];
  const ax = 'xyz'[axis];
  const rel = mc['eps_r_' + ax], sig = mc['sigma_e_' + ax];
  const [nSelf, n1, n2] = COEFF_GROUPS[L.comp];
  const cs = coeffs[nSelf], c1 = coeffs[n1], c2 = coeffs[n2];
  const fi = elementFieldIndices(elementIndices(el, grid), axis, grid);
  const da = d[axis], dt1 = d[L.t1], dt2 = d[L.t2];

  let aTerm = 0, aSign = -1, R = 0;
  if (kind === 'capacitor') {
    aTerm = (2 * el.capacitance_per_component * da) / (dt1 * dt2);
    aSign = +1;
  } else {
    R = el.resistance_per_component;
    if (kind === 'vsource' && R === 0) R = 1e-20;
    aTerm = (dt * da) / (R * dt1 * dt2);
  }
  const srcCoef = kind === 'vsource' ? new Float64Array(fi.length) : null;
  for (let n = 0; n < fi.length; n++) {
    const idx = fi[n];
    const denom = 2 * EPS0 * rel[idx] + dt * sig[idx] + aTerm;
    cs[idx] = aSign < 0
      ? (2 * EPS0 * rel[idx] - dt * sig[idx] - aTerm) / denom
      : (2 * EPS0 * rel[idx] - dt * sig[idx] + aTerm) / denom;
    c1[idx] = (2 * dt / dt1) / denom;
    c2[idx] = -(2 * dt / dt2) / denom;
    if (srcCoef) srcCoef[n] = -(2 * dt / (R * dt1 * dt2)) / denom;
  }
  el.field_indices = fi;
  if (srcCoef) el[L.src] = srcCoef;
}

/** Per-component resistance / capacitance of a lumped element box (sources.js:196-260). */
export function initPassiveElement(el, grid, kind) {
  const b = elementIndices(el, grid);
  Object.assign(el, b);
  const ext = [1 + b.ie - b.is, 1 + b.je - b.js, 1 + b.ke - b.ks];
  const gap = [b.ie - b.is, b.je - b.js, b.ke - b.ks];
  const axis = AXIS_OF[el.direction[0]];
  const o1 = axis === 0 ? 1 : 0, o2 = axis === 2 ? 1 : 2;
  if (kind === 'resistor') el.resistance_per_component = ext[o1] * ext[o2] / gap[axis] * el.resistance;
  else el.capacitance_per_component = gap[axis] / (ext[o1] * ext[o2]) * el.capacitance;
}

export function applyLumpedElementCoefficients(coeffs, mc, grid, voltageSources, resistors, capacitors) {
  for (const vs of voltageSources) applyLumped(coeffs, mc, grid, vs, 'vsource');
  for (const r of resistors || []) applyLumped(coeffs, mc, grid, r, 'resistor');
  for (const c of capacitors || []) applyLumped(coeffs, mc, grid, c, 'capacitor');
}
