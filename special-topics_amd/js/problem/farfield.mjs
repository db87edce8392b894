// Near-to-far-field (NTFF) Huygens box set-up.
//
// Box bounds and accumulator layout follow js_fdtd/src/farfield.js:28-78: the box
// sits `number_of_cells_from_outer_boundary` cells inside the domain, each of the 24
// tangential J/M components of a face has {re, im} Float64Arrays of nFreq*faceSize
// entries, index f*faceSize + (row*ncol + col).  phaseTable() precomputes the
// per-step DFT kernels cos(w dt (ts + 1/2)), -sin(.), cos(w dt (ts + 1)), -sin(.)
// exactly as accumulateFarfieldDFT evaluates them (farfield.js:101-110), so native
// backends use the identical V8 Math.cos/Math.sin values.

// Accumulator names per face, in the order the reference declares them.
export const FF_ARRAYS = {
  xp: ['cjyxp', 'cjzxp', 'cmyxp', 'cmzxp'], xn: ['cjyxn', 'cjzxn', 'cmyxn', 'cmzxn'],
  yp: ['cjxyp', 'cjzyp', 'cmxyp', 'cmzyp'], yn: ['cjxyn', 'cjzyn', 'cmxyn', 'cmzyn'],
  zp: ['cjxzp', 'cjyzp', 'cmxzp', 'cmyzp'], zn: ['cjxzn', 'cjyzn', 'cmxzn', 'cmyzn'],
};
// Native ABI numbering of the 24 accumulators.
export const FF_ORDER = [].concat(FF_ARRAYS.xp, FF_ARRAYS.xn, FF_ARRAYS.yp, FF_ARRAYS.yn,
                                  FF_ARRAYS.zp, FF_ARRAYS.zn);

export function initFarfield(params, grid) {
  const { nx, ny, nz } = grid;
  const frequencies = params.frequencies || [];
  const nc = params.number_of_cells_from_outer_boundary;
  const nFreq = frequencies.length;
  if (nFreq === 0) return { nFreq: 0 };
  const li = nc + 1, lj = nc + 1, lk = nc + 1;
  const ui = nx - nc + 1, uj = ny - nc + 1, uk = nz - nc + 1;
  const ni_ff = ui - li, nj_ff = uj - lj, nk_ff = uk - lk;
  const sz = { x: nj_ff * nk_ff, y: ni_ff * nk_ff, z: ni_ff * nj_ff };
  const ff = {
    nFreq, li, lj, lk, ui, uj, uk, ni_ff, nj_ff, nk_ff,
    sz_xface: sz.x, sz_yface: sz.y, sz_zface: sz.z,
    frequencies,
    w: Float64Array.from(frequencies, (f) => 2 * Math.PI * f),
  };
  for (const face of Object.keys(FF_ARRAYS)) {
    const n = nFreq * sz[face[0]];
    for (const name of FF_ARRAYS[face]) ff[name] = { re: new Float64Array(n), im: new Float64Array(n) };
  }
  return ff;
}

/** Float64Array [steps][nFreq][4] = cos_h, sin_h, cos_e, sin_e (sin already negated). */
export function phaseTable(ff, dt, steps) {
  const out = new Float64Array(steps * ff.nFreq * 4);
  let p = 0;
  for (let ts = 0; ts < steps; ts++) {
    for (let mi = 0; mi < ff.nFreq; mi++) {
      const wdt = ff.w[mi] * dt;
      const ph = wdt * (ts + 0.5);
      const pe = wdt * (ts + 1);
      out[p++] = Math.cos(ph);
      out[p++] = -Math.sin(ph);
      out[p++] = Math.cos(pe);
      out[p++] = -Math.sin(pe);
    }
  }
  return out;
}
