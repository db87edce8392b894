// CPML set-up: 1-D graded profiles, Psi/CPsi storage per face, and the in-place
// division of the face's curl coefficients by kappa.
//
// Follows js_fdtd/src/cpml.js:25-448.  Per face along axis a (side n or p) two H
// components and two E components carry a convolution term.  Psi and CPsi of a face
// have the component's shape with the extent along a replaced by nc.  CPsi is the
// curl coefficient times the cell size taken BEFORE the kappa division
// (cpml.js:177-210 precede :212-243), and the objects returned carry the
// reference's field names so that the hip backend accepts problems built by either
// this module or js_fdtd's own setup.

import { EPS0 } from './constants.mjs';
import { componentDims } from './grid.mjs';

const MU0_LOCAL = 4 * Math.PI * 1e-7;

// For each axis: [component, curl coefficient, source field of the Psi difference,
// Psi/CPsi name suffix] for the two H and the two E components, in the order the
// reference keeps them.
export const FACE_TERMS = {
  x: {
    h: [['Hy', 'Chyez', 'Ez', 'hyx'], ['Hz', 'Chzey', 'Ey', 'hzx']],
    e: [['Ey', 'Ceyhz', 'Hz', 'eyx'], ['Ez', 'Cezhy', 'Hy', 'ezx']],
  },
  y: {
    h: [['Hx', 'Chxez', 'Ez', 'hxy'], ['Hz', 'Chzex', 'Ex', 'hzy']],
    e: [['Ez', 'Cezhx', 'Hx', 'ezy'], ['Ex', 'Cexhz', 'Hz', 'exy']],
  },
  z: {
    h: [['Hx', 'Chxey', 'Ey', 'hxz'], ['Hy', 'Chyex', 'Ex', 'hyz']],
    e: [['Ex', 'Cexhy', 'Hy', 'exz'], ['Ey', 'Ceyhx', 'Hx', 'eyz']],
  },
};

export const FACE_NAMES = ['xn', 'xp', 'yn', 'yp', 'zn', 'zp'];

/** Graded b/a/kappa profiles of one face, index 0 = the cell next to the interior
 *  for an ascending (p) face and next to the outer wall for a descending (n) one. */
export function cpmlProfile(nc, cellSize, dt, order, sigmaFactor, kappaMax, alphaMin, alphaMax, ascending) {
  const sigmaMax = sigmaFactor * (order + 1) / (150 * Math.PI * cellSize);
  const prof = {
    b_e: new Float64Array(nc), a_e: new Float64Array(nc),
    b_m: new Float64Array(nc), a_m: new Float64Array(nc),
    kappa_e: new Float64Array(nc), kappa_m: new Float64Array(nc),
  };
  for (let ci = 0; ci < nc; ci++) {
    const rhoE = ascending ? (ci + 1 - 0.75) / nc : (nc - ci - 0.75) / nc;
    const rhoM = ascending ? (ci + 1 - 0.25) / nc : (nc - ci - 0.25) / nc;
    const se = sigmaMax * Math.pow(rhoE, order);
    const sm = (MU0_LOCAL / EPS0) * sigmaMax * Math.pow(rhoM, order);
    const ke = 1 + (kappaMax - 1) * Math.pow(rhoE, order);
    const km = 1 + (kappaMax - 1) * Math.pow(rhoM, order);
    const ae = alphaMin + (alphaMax - alphaMin) * (1 - rhoE);
    const am = (MU0_LOCAL / EPS0) * (alphaMin + (alphaMax - alphaMin) * (1 - rhoM));
    const be = Math.exp(-(dt / EPS0) * (se / ke + ae));
    const bm = Math.exp(-(dt / MU0_LOCAL) * (sm / km + am));
    prof.b_e[ci] = be;
    prof.a_e[ci] = (1 / cellSize) * (be - 1) * se / (ke * (se + ke * ae));
    prof.b_m[ci] = bm;
    prof.a_m[ci] = (1 / cellSize) * (bm - 1) * sm / (km * (sm + km * am));
    prof.kappa_e[ci] = ke;
    prof.kappa_m[ci] = km;
  }
  return prof;
}

/** Plane index (along the face axis) that Psi slot ci of a face acts on. */
export function facePlanes(n, nc, ascending) {
  return {
    mStart: ascending ? n - nc : 0,      // H planes mStart .. mStart+nc-1
    eStart: ascending ? n - nc : 1,      // E planes eStart .. eStart+nc-1
  };
}

// Copy plane `plane` (along `axis`) of a component array into slot ci of a face
// array (or the reverse when `into` is false), applying fn to each element pair.
function forPlane(dims, axis, plane, nc, ci, fn) {
  const fdims = dims.slice();
  fdims[axis] = nc;
  const lo = [0, 0, 0], hi = dims.slice();
  lo[axis] = plane; hi[axis] = plane + 1;
  for (let i = lo[0]; i < hi[0]; i++) {
    for (let j = lo[1]; j < hi[1]; j++) {
      for (let k = lo[2]; k < hi[2]; k++) {
        const fi = axis === 0 ? ci : i, fj = axis === 1 ? ci : j, fk = axis === 2 ? ci : k;
        fn((i * dims[1] + j) * dims[2] + k, (fi * fdims[1] + fj) * fdims[2] + fk);
      }
    }
  }
}

function initFace(face, nc, grid, params, coeffs) {
  const { nx, ny, nz, dt } = grid;
  const axisName = face[0];
  const axis = 'xyz'.indexOf(axisName);
  const n = [nx, ny, nz][axis];
  const cell = [grid.dx, grid.dy, grid.dz][axis];
  const ascending = face[1] === 'p';
  const prof = cpmlProfile(nc, cell, dt, params.order, params.sigmaFactor, params.kappaMax,
                           params.alphaMin, params.alphaMax, ascending);
  const { mStart, eStart } = facePlanes(n, nc, ascending);
  const out = { face, nc, b_e: prof.b_e, a_e: prof.a_e, b_m: prof.b_m, a_m: prof.a_m,
                ascending, nx, ny, nz, nxp1: nx + 1, nyp1: ny + 1, nzp1: nz + 1 };
  const terms = FACE_TERMS[axisName];

  // Storage and CPsi = coefficient * cell size, all taken before any kappa division.
  const groups = [[terms.h, mStart], [terms.e, eStart]];
  for (const [list, start] of groups) {
    for (const [comp, coef, , suffix] of list) {
      const dims = componentDims(comp, nx, ny, nz);
      const len = (dims[0] * dims[1] * dims[2] / dims[axis]) * nc;
      out['Psi_' + suffix] = new Float64Array(len);
      const cpsi = new Float64Array(len);
      const c = coeffs[coef];
      for (let ci = 0; ci < nc; ci++) {
        const plane = start + ci;
        if (plane < 0 || plane >= dims[axis]) continue;
        forPlane(dims, axis, plane, nc, ci, (src, dst) => { cpsi[dst] = c[src] * cell; });
      }
      out['CPsi_' + suffix] = cpsi;
    }
  }
  // Kappa division of the curl coefficients that the convolution term replaces.
  for (let ci = 0; ci < nc; ci++) {
    for (const [list, start, kap] of [[terms.h, mStart, prof.kappa_m], [terms.e, eStart, prof.kappa_e]]) {
      for (const [comp, coef] of list) {
        const dims = componentDims(comp, nx, ny, nz);
        const plane = start + ci;
        if (plane < 0 || plane >= dims[axis]) continue;
        const c = coeffs[coef];
        const kv = kap[ci];
        forPlane(dims, axis, plane, nc, ci, (src) => { c[src] /= kv; });
      }
    }
  }
  out[axisName.replace('x', 'i').replace('y', 'j').replace('z', 'k') + '_m_start'] = mStart;
  out[axisName.replace('x', 'i').replace('y', 'j').replace('z', 'k') + '_e_start'] = eStart;
  return out;
}

/** All CPML faces of a boundary description; modifies coeffs in place. */
export function initCPML(boundary, coeffs, grid) {
  const params = {
    order: boundary.cpml_order, sigmaFactor: boundary.cpml_sigma_factor,
    kappaMax: boundary.cpml_kappa_max, alphaMin: boundary.cpml_alpha_min,
    alphaMax: boundary.cpml_alpha_max,
  };
  const cpml = {};
  for (const face of FACE_NAMES) {
    if (boundary['type_' + face] === 'cpml') {
      cpml[face] = initFace(face, boundary['cpml_cells_' + face], grid, params, coeffs);
    }
  }
  return cpml;
}
