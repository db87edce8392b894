// Port voltage / current observers and point field probes.
//
// Initialisation follows js_fdtd/src/sampling.js:51-136.  currentLoopTerms() turns a
// current observer into the ordered list of (H component, flat index, weight) terms
// whose left-to-right sum is the Ampere loop of captureSampledCurrents
// (sampling.js:204-245): four segments, +d*H on the first two, -d*H on the last two,
// then a sign flip for 'n' directions.  Native backends sum that list in order, which
// reproduces the reference's floating-point result exactly (a - b*H == a + (-b)*H).

import { boxIndexList } from './grid.mjs';

function coordIdx(coord, domainMin, cellSize) {
  return Math.round((coord - domainMin) / cellSize) + 1;
}

function timeAxis(n, dt, offset) {
  const t = new Float64Array(n);
  for (let i = 0; i < n; i++) t[i] = (i + offset) * dt;
  return t;
}

export function initSampledVoltages(list, grid) {
  const { min_x, min_y, min_z, dx, dy, dz, ny, nz, nyp1, nzp1, numberOfTimeSteps, dt } = grid;
  for (const obs of list) {
    const is = coordIdx(obs.min_x, min_x, dx), js = coordIdx(obs.min_y, min_y, dy);
    const ks = coordIdx(obs.min_z, min_z, dz), ie = coordIdx(obs.max_x, min_x, dx);
    const je = coordIdx(obs.max_y, min_y, dy), ke = coordIdx(obs.max_z, min_z, dz);
    Object.assign(obs, { is, js, ks, ie, je, ke });
    const dir = obs.direction[0];
    if (dir === 'x') {
      obs.field_indices = boxIndexList(is, ie - 1, js, je, ks, ke, nyp1, nzp1);
      obs.Csvf = -dx / ((je - js + 1) * (ke - ks + 1));
    } else if (dir === 'y') {
      obs.field_indices = boxIndexList(is, ie, js, je - 1, ks, ke, ny, nzp1);
      obs.Csvf = -dy / ((ke - ks + 1) * (ie - is + 1));
    } else {
      obs.field_indices = boxIndexList(is, ie, js, je, ks, ke - 1, nyp1, nz);
      obs.Csvf = -dz / ((ie - is + 1) * (je - js + 1));
    }
    if (obs.direction[1] === 'n') obs.Csvf *= -1;
    obs.sampled_value = new Float64Array(numberOfTimeSteps);
    obs.time = timeAxis(numberOfTimeSteps, dt, 1);
  }
}

export function initSampledCurrents(list, grid) {
  const { min_x, min_y, min_z, dx, dy, dz, numberOfTimeSteps, dt } = grid;
  for (const obs of list) {
    obs.is = coordIdx(obs.min_x, min_x, dx); obs.js = coordIdx(obs.min_y, min_y, dy);
    obs.ks = coordIdx(obs.min_z, min_z, dz); obs.ie = coordIdx(obs.max_x, min_x, dx);
    obs.je = coordIdx(obs.max_y, min_y, dy); obs.ke = coordIdx(obs.max_z, min_z, dz);
    obs.sampled_value = new Float64Array(numberOfTimeSteps);
    obs.time = timeAxis(numberOfTimeSteps, dt, 0.5);
  }
}

export function initSampledElectricFields(list, grid) {
  const { min_x, min_y, min_z, dx, dy, dz, numberOfTimeSteps, dt } = grid;
  for (const obs of list) {
    obs.is = coordIdx(obs.x, min_x, dx); obs.js = coordIdx(obs.y, min_y, dy);
    obs.ks = coordIdx(obs.z, min_z, dz);
    obs.sampled_value = new Float64Array(numberOfTimeSteps);
    obs.time = timeAxis(numberOfTimeSteps, dt, 1);
  }
}

export function initSampledMagneticFields(list, grid) {
  const { min_x, min_y, min_z, dx, dy, dz, numberOfTimeSteps, dt } = grid;
  for (const obs of list) {
    obs.is = coordIdx(obs.x, min_x, dx); obs.js = coordIdx(obs.y, min_y, dy);
    obs.ks = coordIdx(obs.z, min_z, dz);
    obs.sampled_value = new Float64Array(numberOfTimeSteps);
    obs.time = timeAxis(numberOfTimeSteps, dt, 0.5);
  }
}

// H component ids used by the native ABI (0 = Hx, 1 = Hy, 2 = Hz).
const HX = 0, HY = 1, HZ = 2;

/**
 * Ordered Ampere-loop terms of a current observer.  Returns
 * { comp: Int32Array, index: Int32Array, weight: Float64Array, negate: 0|1 } with
 * flat indices into the H component arrays (reference layout).  Throws if a term
 * falls outside its array (the reference would read `undefined` and record NaN).
 */
export function currentLoopTerms(obs, grid) {
  const { dx, dy, dz, nx, ny, nz, nyp1, nzp1 } = grid;
  const I = obs.is - 1, J = obs.js - 1, K = obs.ks - 1;
  const IE = obs.ie - 1, JE = obs.je - 1, KE = obs.ke - 1;
  const comp = [], index = [], weight = [];
  const dims = [[nx + 1, ny, nz], [nx, nyp1, nz], [nx, ny, nzp1]];
  const push = (c, i, j, k, w) => {
    const d = dims[c];
    if (i < 0 || j < 0 || k < 0 || i >= d[0] || j >= d[1] || k >= d[2]) {
      throw new Error(`current observer ${obs.label || ''}: H term (${c},${i},${j},${k}) outside the grid`);
    }
    comp.push(c); index.push((i * d[1] + j) * d[2] + k); weight.push(w);
  };
  const range = (a, b, fn) => { for (let t = a; t <= b; t++) fn(t); };
  switch (obs.direction[0]) {
    case 'x': {
      const ix = IE - 1;
      range(J, JE - 1, (j) => push(HY, ix, j, K - 1, dy));
      range(K, KE - 1, (k) => push(HZ, ix, JE - 1, k, dz));
      range(J, JE - 1, (j) => push(HY, ix, j, KE - 1, -dy));
      range(K, KE - 1, (k) => push(HZ, ix, J - 1, k, -dz));
      break;
    }
    case 'y': {
      const jx = JE - 1;
      range(K, KE - 1, (k) => push(HZ, I - 1, jx, k, dz));
      range(I, IE - 1, (i) => push(HX, i, jx, KE - 1, dx));
      range(K, KE - 1, (k) => push(HZ, IE, jx, k, -dz));
      range(I, IE - 1, (i) => push(HX, i, jx, K - 1, -dx));
      break;
    }
    default: {
      const kx = KE - 1;
      range(I, IE - 1, (i) => push(HX, i, J - 1, kx, dx));
      range(J, JE - 1, (j) => push(HY, IE, j, kx, dy));
      range(I, IE - 1, (i) => push(HX, i, JE, kx, -dx));
      range(J, JE - 1, (j) => push(HY, I - 1, j, kx, -dy));
    }
  }
  return {
    comp: Int32Array.from(comp), index: Int32Array.from(index),
    weight: Float64Array.from(weight), negate: obs.direction[1] === 'n' ? 1 : 0,
  };
}
