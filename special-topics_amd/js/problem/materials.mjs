// Material index grid and the per-component material averages.
//
// Semantics follow js_fdtd/src/materials.js: the Uint8 cell grid of
// buildMaterialGrid (:21-68, spheres then bricks, 1-based material ids), the Yee
// averaging of computeMaterialComponents (:87-238: arithmetic 4-cell mean of eps and
// sigma_e on E edges, harmonic 2-cell mean of mu and sigma_m on H faces) and the
// zero-thickness conductor plates of applyPECPlates (:253-332).  The implementation
// here is table driven over the three axes instead of one block per component; the
// floating-point expression trees are kept identical because the coefficient arrays
// must match the reference bit for bit.

import { componentDims } from './grid.mjs';

const AXES = ['x', 'y', 'z'];

/** Uint8 material ids per cell, flat (i*ny + j)*nz + k, default 1 (= first type). */
export function buildMaterialGrid(grid, bricks, spheres) {
  const { nx, ny, nz, dx, dy, dz, min_x, min_y, min_z } = grid;
  const mat = new Uint8Array(nx * ny * nz).fill(1);

  for (const s of spheres) {
    const r2 = s.radius * s.radius;
    for (let i = 0; i < nx; i++) {
      const ex = (i + 0.5) * dx + min_x - s.center_x;
      for (let j = 0; j < ny; j++) {
        const ey = (j + 0.5) * dy + min_y - s.center_y;
        for (let k = 0; k < nz; k++) {
          const ez = (k + 0.5) * dz + min_z - s.center_z;
          if (ex * ex + ey * ey + ez * ez <= r2) mat[(i * ny + j) * nz + k] = s.material_type;
        }
      }
    }
  }

  for (const b of bricks) {
    // 0-based half-open cell ranges [lo, hi) clipped to the grid.
    const lo = [Math.round((b.min_x - min_x) / dx), Math.round((b.min_y - min_y) / dy),
                Math.round((b.min_z - min_z) / dz)];
    const hi = [Math.round((b.max_x - min_x) / dx), Math.round((b.max_y - min_y) / dy),
                Math.round((b.max_z - min_z) / dz)];
    const i0 = Math.max(lo[0], 0), i1 = Math.min(hi[0], nx);
    const j0 = Math.max(lo[1], 0), j1 = Math.min(hi[1], ny);
    const k0 = Math.max(lo[2], 0), k1 = Math.min(hi[2], nz);
    for (let i = i0; i < i1; i++) {
      for (let j = j0; j < j1; j++) {
        const row = (i * ny + j) * nz;
        for (let k = k0; k < k1; k++) mat[row + k] = b.material_type;
      }
    }
  }
  return mat;
}

function propertyTables(materialTypes) {
  const n = materialTypes.length + 1;
  const t = { eps: new Float64Array(n), mu: new Float64Array(n),
              sige: new Float64Array(n), sigm: new Float64Array(n) };
  materialTypes.forEach((m, ind) => {
    t.eps[ind + 1] = m.eps_r;
    t.mu[ind + 1] = m.mu_r || 1e-20;       // a zero mu or sigma_m would divide by zero
    t.sige[ind + 1] = m.sigma_e;
    t.sigm[ind + 1] = m.sigma_m || 1e-20;
  });
  return t;
}

/**
 * eps_r_{x,y,z}, sigma_e_{x,y,z} on E edges and mu_r_{x,y,z}, sigma_m_{x,y,z} on H
 * faces.  Unvisited entries keep the free-space defaults (eps = mu = 1, sigma = 0).
 */
export function computeMaterialComponents(mat, materialTypes, grid) {
  const { nx, ny, nz } = grid;
  const n = [nx, ny, nz];
  const t = propertyTables(materialTypes);
  const cellStride = [ny * nz, nz, 1];
  const out = {};

  for (let c = 0; c < 3; c++) {
    const ax = AXES[c];
    // E edge along axis c: mean of the 4 cells sharing it.  The two transverse axes
    // are taken in x,y,z order; summation order is p, p-e1, p-e2, p-e1-e2.
    {
      const dims = componentDims('E' + ax, nx, ny, nz);
      const eps = new Float64Array(dims[0] * dims[1] * dims[2]).fill(1.0);
      const sig = new Float64Array(dims[0] * dims[1] * dims[2]);
      const t1 = c === 0 ? 1 : 0;
      const t2 = c === 2 ? 1 : 2;
      const s1 = cellStride[t1], s2 = cellStride[t2];
      const lo = [1, 1, 1];
      lo[c] = 0;
      for (let i = lo[0]; i < n[0]; i++) {
        for (let j = lo[1]; j < n[1]; j++) {
          for (let k = lo[2]; k < n[2]; k++) {
            const p = (i * ny + j) * nz + k;
            const m1 = mat[p], m2 = mat[p - s1], m3 = mat[p - s2], m4 = mat[p - s1 - s2];
            const idx = (i * dims[1] + j) * dims[2] + k;
            eps[idx] = 0.25 * (t.eps[m1] + t.eps[m2] + t.eps[m3] + t.eps[m4]);
            sig[idx] = 0.25 * (t.sige[m1] + t.sige[m2] + t.sige[m3] + t.sige[m4]);
          }
        }
      }
      out['eps_r_' + ax] = eps;
      out['sigma_e_' + ax] = sig;
    }
    // H face normal to axis c: harmonic mean of the 2 cells sharing it.
    {
      const dims = componentDims('H' + ax, nx, ny, nz);
      const mu = new Float64Array(dims[0] * dims[1] * dims[2]).fill(1.0);
      const sig = new Float64Array(dims[0] * dims[1] * dims[2]);
      const s = cellStride[c];
      const lo = [0, 0, 0];
      lo[c] = 1;
      for (let i = lo[0]; i < n[0]; i++) {
        for (let j = lo[1]; j < n[1]; j++) {
          for (let k = lo[2]; k < n[2]; k++) {
            const p = (i * ny + j) * nz + k;
            const idx = (i * dims[1] + j) * dims[2] + k;
            const mu1 = t.mu[mat[p]], mu2 = t.mu[mat[p - s]];
            mu[idx] = 2 * mu1 * mu2 / (mu1 + mu2);
            const sm1 = t.sigm[mat[p]], sm2 = t.sigm[mat[p - s]];
            sig[idx] = 2 * sm1 * sm2 / (sm1 + sm2);
          }
        }
      }
      out['mu_r_' + ax] = mu;
      out['sigma_m_' + ax] = sig;
    }
  }
  return out;
}

/**
 * Zero-thickness conducting bricks: a brick flat along axis a sets sigma_e on the E
 * edges lying in that plane to the material's sigma_e (materials.js:253-332).  Along
 * the edge's own axis the range is the brick's edges [lo-1, hi-1), along the other
 * in-plane axis its nodes [lo-1, hi-1]; out-of-range entries are skipped.
 */
export function applyPECPlates(bricks, materialTypes, mc, grid) {
  const { nx, ny, nz, dx, dy, dz, min_x, min_y, min_z } = grid;
  for (const b of bricks) {
    const sigmaPec = materialTypes[b.material_type - 1].sigma_e;
    if (sigmaPec === 0) continue;
    const lo = [Math.round((b.min_x - min_x) / dx) + 1, Math.round((b.min_y - min_y) / dy) + 1,
                Math.round((b.min_z - min_z) / dz) + 1];
    const hi = [Math.round((b.max_x - min_x) / dx) + 1, Math.round((b.max_y - min_y) / dy) + 1,
                Math.round((b.max_z - min_z) / dz) + 1];
    for (let a = 0; a < 3; a++) {
      if (lo[a] !== hi[a]) continue;
      for (let c = 0; c < 3; c++) {
        if (c === a) continue;
        const dims = componentDims('E' + AXES[c], nx, ny, nz);
        const arr = mc['sigma_e_' + AXES[c]];
        const r0 = [0, 0, 0], r1 = [0, 0, 0];
        for (let q = 0; q < 3; q++) {
          if (q === a) { r0[q] = lo[q] - 1; r1[q] = lo[q]; }
          else if (q === c) { r0[q] = lo[q] - 1; r1[q] = hi[q] - 1; }
          else { r0[q] = lo[q] - 1; r1[q] = hi[q]; }
        }
        for (let i = r0[0]; i < r1[0]; i++) {
          if (i < 0 || i >= dims[0]) continue;
          for (let j = r0[1]; j < r1[1]; j++) {
            if (j < 0 || j >= dims[1]) continue;
            for (let k = r0[2]; k < r1[2]; k++) {
              if (k < 0 || k >= dims[2]) continue;
              arr[(i * dims[1] + j) * dims[2] + k] = sigmaPec;
            }
          }
        }
      }
    }
  }
}
