// Near-to-far-field post-processing of the Huygens-box DFT accumulators, FP64.
//
// calcRadiatedPower: P(f) = 0.5 * sum over the six faces of +-dA * Re(M x conj(J)).n
// (js_fdtd/src/farfield.js:257-314; face order zp, zn, yp, yn, xp, xn).
// calcDirectivity: radiation vectors N (from J) and L (from M) summed over every
// surface cell with the phase k r^.r', then
//   D_theta = k^2/(8 pi eta0 P) |L_phi + eta0 N_theta|^2,
//   D_phi   = k^2/(8 pi eta0 P) |L_theta - eta0 N_phi|^2
// (farfield.js:328-607).  The per-cell projections, phases and accumulation order of
// every sum follow the reference term by term, so the results agree bit for bit under
// the same Math library.  The loops run angle-outer (one scalar accumulator per angle),
// which leaves every accumulator's sequence of additions unchanged.

import { C0, ETA0 } from '../problem/constants.mjs';

const FACE_KIND = {
  // Tangential J and M components of each face pair, the surface axes, and whether
  // N_phi has one term (x and y faces) or two (z faces).
  x: { J: ['cjy', 'cjz'], M: ['cmy', 'cmz'], rows: 'nj_ff', cols: 'nk_ff', size: 'sz_xface', npTerms: 1 },
  y: { J: ['cjx', 'cjz'], M: ['cmx', 'cmz'], rows: 'ni_ff', cols: 'nk_ff', size: 'sz_yface', npTerms: 1 },
  z: { J: ['cjx', 'cjy'], M: ['cmx', 'cmy'], rows: 'ni_ff', cols: 'nj_ff', size: 'sz_zface', npTerms: 2 },
};

export function calcRadiatedPower(ff, grid) {
  const { dx, dy, dz } = grid;
  const P = new Float64Array(ff.nFreq);
  // Re(A conj(B)) - Re(C conj(D)) over one face, added with sign s and area dA.
  const face = (acc, s, dA, A, B, C, D, o, n) => {
    for (let p = o; p < o + n; p++) {
      const t = dA * (A.re[p] * B.re[p] + A.im[p] * B.im[p] - C.re[p] * D.re[p] - C.im[p] * D.im[p]);
      acc = s > 0 ? acc + t : acc - t;
    }
    return acc;
  };
  for (let m = 0; m < ff.nFreq; m++) {
    const ox = m * ff.sz_xface, oy = m * ff.sz_yface, oz = m * ff.sz_zface;
    let acc = 0;
    acc = face(acc, +1, dx * dy, ff.cmyzp, ff.cjxzp, ff.cmxzp, ff.cjyzp, oz, ff.sz_zface);
    acc = face(acc, -1, dx * dy, ff.cmyzn, ff.cjxzn, ff.cmxzn, ff.cjyzn, oz, ff.sz_zface);
    acc = face(acc, +1, dx * dz, ff.cmxyp, ff.cjzyp, ff.cmzyp, ff.cjxyp, oy, ff.sz_yface);
    acc = face(acc, -1, dx * dz, ff.cmxyn, ff.cjzyn, ff.cmzyn, ff.cjxyn, oy, ff.sz_yface);
    acc = face(acc, +1, dy * dz, ff.cmzxp, ff.cjyxp, ff.cmyxp, ff.cjzxp, ox, ff.sz_xface);
    acc = face(acc, -1, dy * dz, ff.cmzxn, ff.cjyxn, ff.cmyxn, ff.cjzxn, ox, ff.sz_xface);
    P[m] = 0.5 * acc;
  }
  return P;
}

/**
 * Directivity over (theta[a], phi[a]) cuts at every NTFF frequency.
 * @returns {{dataTheta: Float64Array[], dataPhi: Float64Array[]}} per frequency.
 */
export function calcDirectivity(ff, grid, radPower, theta, phi) {
  const { dx, dy, dz } = grid;
  const { nFreq, li, lj, lk, ui, uj, uk } = ff;
  const nA = theta.length;
  const dataTheta = [], dataPhi = [];
  const sinth = Float64Array.from(theta, Math.sin), costh = Float64Array.from(theta, Math.cos);
  const sinph = Float64Array.from(phi, Math.sin), cosph = Float64Array.from(phi, Math.cos);
  const ci = 0.5 * (ui + li), cj = 0.5 * (uj + lj), ck = 0.5 * (uk + lk);
  const area = { x: dy * dz, y: dx * dz, z: dx * dy };
  // Offsets of the two faces of each pair along their normal, and of cell centres.
  const normalOff = { x: [ui - ci, li - ci], y: [uj - cj, lj - cj], z: [uk - ck, lk - ck] };
  const cellOff = {
    i: (n) => (li - 1) + n - ci + 0.5, j: (n) => (lj - 1) + n - cj + 0.5, k: (n) => (lk - 1) + n - ck + 0.5,
  };

  for (let m = 0; m < nFreq; m++) {
    const k = 2 * Math.PI * ff.frequencies[m] / C0;
    const scale = k * k / (8 * Math.PI * ETA0 * radPower[m]);
    const Dt = new Float64Array(nA), Dp = new Float64Array(nA);
    for (let a = 0; a < nA; a++) {
      const st = sinth[a], ct = costh[a], sp = sinph[a], cp = cosph[a];
      // r^.r' as the reference evaluates it: x term + y term + z term.
      const rdot = (ox, oy, oz) => ox * dx * st * cp + oy * dy * st * sp + oz * dz * ct;
      let NtR = 0, NtI = 0, NpR = 0, NpI = 0, LtR = 0, LtI = 0, LpR = 0, LpI = 0;
      for (const axis of ['x', 'y', 'z']) {
        const K = FACE_KIND[axis];
        const dA = area[axis];
        // N_theta = J1*A1 + J2*A2, N_phi = J1*B1 (+ J2*B2 on z faces); L likewise from M.
        let A1, A2, B1, B2 = 0;
        if (axis === 'x') { A1 = dA * (ct * sp); A2 = -(dA * st); B1 = dA * cp; }
        else if (axis === 'y') { A1 = dA * (ct * cp); A2 = -(dA * st); B1 = -(dA * sp); }
        else { A1 = dA * (ct * cp); A2 = dA * (ct * sp); B1 = -(dA * sp); B2 = dA * cp; }
        const two = K.npTerms === 2;
        const rows = ff[K.rows], cols = ff[K.cols], o = m * ff[K.size];
        const sides = [
          { J1: ff[K.J[0] + axis + 'p'], J2: ff[K.J[1] + axis + 'p'], M1: ff[K.M[0] + axis + 'p'],
            M2: ff[K.M[1] + axis + 'p'], off: normalOff[axis][0] },
          { J1: ff[K.J[0] + axis + 'n'], J2: ff[K.J[1] + axis + 'n'], M1: ff[K.M[0] + axis + 'n'],
            M2: ff[K.M[1] + axis + 'n'], off: normalOff[axis][1] },
        ];
        for (let r = 0; r < rows; r++) {
          for (let c = 0; c < cols; c++) {
            const p = o + r * cols + c;
            for (const S of sides) {
              let ph;
              if (axis === 'x') ph = rdot(S.off, cellOff.j(r), cellOff.k(c));
              else if (axis === 'y') ph = rdot(cellOff.i(r), S.off, cellOff.k(c));
              else ph = rdot(cellOff.i(r), cellOff.j(c), S.off);
              const cs = Math.cos(k * ph), sn = Math.sin(k * ph);
              const j1r = S.J1.re[p], j1i = S.J1.im[p], j2r = S.J2.re[p], j2i = S.J2.im[p];
              const m1r = S.M1.re[p], m1i = S.M1.im[p], m2r = S.M2.re[p], m2i = S.M2.im[p];
              let xr = j1r * A1 + j2r * A2, xi = j1i * A1 + j2i * A2;
              NtR += xr * cs - xi * sn; NtI += xr * sn + xi * cs;
              xr = two ? j1r * B1 + j2r * B2 : j1r * B1; xi = two ? j1i * B1 + j2i * B2 : j1i * B1;
              NpR += xr * cs - xi * sn; NpI += xr * sn + xi * cs;
              xr = m1r * A1 + m2r * A2; xi = m1i * A1 + m2i * A2;
              LtR += xr * cs - xi * sn; LtI += xr * sn + xi * cs;
              xr = two ? m1r * B1 + m2r * B2 : m1r * B1; xi = two ? m1i * B1 + m2i * B2 : m1i * B1;
              LpR += xr * cs - xi * sn; LpI += xr * sn + xi * cs;
            }
          }
        }
      }
      const tr = LpR + ETA0 * NtR, ti = LpI + ETA0 * NtI;
      const pr = LtR - ETA0 * NpR, pi = LtI - ETA0 * NpI;
      Dt[a] = scale * (tr * tr + ti * ti);
      Dp[a] = scale * (pr * pr + pi * pi);
    }
    dataTheta.push(Dt);
    dataPhi.push(Dp);
  }
  return { dataTheta, dataPhi };
}

/** The xy (theta = 90 deg) and xz cuts over 0..359 deg that js_fdtd/src/index.js:77-92 writes. */
export function standardCuts() {
  const angles = new Float64Array(360);
  for (let i = 0; i < 360; i++) angles[i] = i * Math.PI / 180;
  const thXY = new Float64Array(360).fill(Math.PI / 2);
  const thXZ = new Float64Array(360), phXZ = new Float64Array(360);
  for (let i = 0; i < 360; i++) {
    const th = angles[i];
    phXZ[i] = th > Math.PI ? Math.PI : 0;
    thXZ[i] = th > Math.PI ? 2 * Math.PI - th : th;
  }
  return { angles, xy: { theta: thXY, phi: angles }, xz: { theta: thXZ, phi: phXZ } };
}
