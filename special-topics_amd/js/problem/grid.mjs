// Domain sizing, time step and the Yee component layout.
//
// Follows js_fdtd/src/grid.js:16-70 (bounding box of all objects, air buffers,
// CPML layers, Math.round snapping, Courant time step) and the component shapes of
// js_fdtd/src/backends/CONTRACT.md:21-41.  Unlike the reference, arrays are plain
// per-component Float64Arrays: the contiguous WebAssembly.Memory of grid.js:125-162
// is a property of the WASM engine, and the HIP backend reproduces the one place
// where that contiguity is observable (the i = nx plane of updateE) explicitly.

import { C0 } from './constants.mjs';

// Field component name -> extent along (i, j, k) as offsets from (nx, ny, nz).
// A "1" means the component has n+1 samples along that axis.
export const COMPONENT_SHAPE = {
  Hx: [1, 0, 0],
  Hy: [0, 1, 0],
  Hz: [0, 0, 1],
  Ex: [0, 1, 1],
  Ey: [1, 0, 1],
  Ez: [1, 1, 0],
};

// The 18 update-coefficient arrays, grouped by the component they update.  The
// first name of each triple multiplies the component itself, the other two the
// two curl differences, in the order the bulk kernels apply them
// (js_fdtd/assembly/fdtd_kernels.ts:54-56,80-82,107-109,150-152,176-178,203-205).
export const COEFF_GROUPS = {
  Hx: ['Chxh', 'Chxey', 'Chxez'],
  Hy: ['Chyh', 'Chyez', 'Chyex'],
  Hz: ['Chzh', 'Chzex', 'Chzey'],
  Ex: ['Cexe', 'Cexhz', 'Cexhy'],
  Ey: ['Ceye', 'Ceyhx', 'Ceyhz'],
  Ez: ['Ceze', 'Cezhy', 'Cezhx'],
};

// Order in which the reference packs the 24 arrays into one buffer
// (js_fdtd/src/grid.js:136-162).  Kept as data because the native ABI numbers the
// coefficient arrays in this order.
export const FIELD_ORDER = ['Hx', 'Hy', 'Hz', 'Ex', 'Ey', 'Ez'];
export const COEFF_ORDER = [
  'Cexe', 'Cexhz', 'Cexhy', 'Ceye', 'Ceyhx', 'Ceyhz', 'Ceze', 'Cezhy', 'Cezhx',
  'Chxh', 'Chxez', 'Chxey', 'Chyh', 'Chyex', 'Chyez', 'Chzh', 'Chzey', 'Chzex',
];

export function componentDims(comp, nx, ny, nz) {
  const s = COMPONENT_SHAPE[comp];
  return [nx + s[0], ny + s[1], nz + s[2]];
}

export function componentLength(comp, nx, ny, nz) {
  const d = componentDims(comp, nx, ny, nz);
  return d[0] * d[1] * d[2];
}

/** Coefficient name -> the field component it belongs to. */
export function coeffComponent(name) {
  for (const comp of FIELD_ORDER) {
    if (COEFF_GROUPS[comp].indexOf(name) >= 0) return comp;
  }
  throw new Error(`unknown coefficient array ${name}`);
}

/**
 * Size the domain and derive dt.  Returns the grid description (no field storage;
 * see allocateFields).  Arithmetic order matches grid.js:21-70 exactly.
 */
export function sizeDomain(params) {
  const { dx, dy, dz, boundary, courantFactor } = params;
  const bricks = params.bricks || [];
  const spheres = params.spheres || [];

  let lo = [Infinity, Infinity, Infinity];
  let hi = [-Infinity, -Infinity, -Infinity];
  const grow = (a0, a1, a2, b0, b1, b2) => {
    lo = [Math.min(lo[0], a0), Math.min(lo[1], a1), Math.min(lo[2], a2)];
    hi = [Math.max(hi[0], b0), Math.max(hi[1], b1), Math.max(hi[2], b2)];
  };
  for (const s of spheres) {
    grow(s.center_x - s.radius, s.center_y - s.radius, s.center_z - s.radius,
         s.center_x + s.radius, s.center_y + s.radius, s.center_z + s.radius);
  }
  for (const b of bricks) grow(b.min_x, b.min_y, b.min_z, b.max_x, b.max_y, b.max_z);

  const cell = [dx, dy, dz];
  const axes = ['x', 'y', 'z'];
  for (let a = 0; a < 3; a++) {
    lo[a] -= cell[a] * boundary['air_buffer_' + axes[a] + 'n'];
    hi[a] += cell[a] * boundary['air_buffer_' + axes[a] + 'p'];
  }
  for (let a = 0; a < 3; a++) {
    for (const side of ['n', 'p']) {
      const f = axes[a] + side;
      const ncells = boundary['cpml_cells_' + f];
      if (boundary['type_' + f] === 'cpml' && ncells > 0) {
        if (side === 'n') lo[a] -= cell[a] * ncells;
        else hi[a] += cell[a] * ncells;
      }
    }
  }

  const n = [0, 1, 2].map((a) => Math.round((hi[a] - lo[a]) / cell[a]));
  const dt = courantFactor / (C0 * Math.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2));

  const [nx, ny, nz] = n;
  return {
    dx, dy, dz, dt,
    nx, ny, nz, nxp1: nx + 1, nyp1: ny + 1, nzp1: nz + 1,
    min_x: lo[0], min_y: lo[1], min_z: lo[2],
    max_x: lo[0] + nx * dx, max_y: lo[1] + ny * dy, max_z: lo[2] + nz * dz,
    numberOfTimeSteps: params.numberOfTimeSteps,
  };
}

/** Zero-initialised field storage, one Float64Array per component. */
export function allocateFields(grid) {
  const fields = {};
  for (const comp of FIELD_ORDER) {
    fields[comp] = new Float64Array(componentLength(comp, grid.nx, grid.ny, grid.nz));
  }
  return fields;
}

/** 1-based grid index of a coordinate: Math.round ties toward +inf (grid.js:249-251). */
export function coordToIndex(coord, domainMin, cellSize) {
  return Math.round((coord - domainMin) / cellSize) + 1;
}

/**
 * Flat 0-based indices of the 1-based inclusive box [is..ie]x[js..je]x[ks..ke] in an
 * array with trailing dims (d2, d3); k outermost, i innermost (grid.js:228-243).
 */
export function boxIndexList(is, ie, js, je, ks, ke, d2, d3) {
  const ni = ie - is + 1, nj = je - js + 1, nk = ke - ks + 1;
  const out = new Int32Array(Math.max(0, ni * nj * nk));
  let p = 0;
  for (let k = ks; k <= ke; k++) {
    for (let j = js; j <= je; j++) {
      for (let i = is; i <= ie; i++) out[p++] = (i - 1) * d2 * d3 + (j - 1) * d3 + (k - 1);
    }
  }
  return out;
}
