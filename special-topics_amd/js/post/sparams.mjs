// Port power waves and the S-matrix, FP64, after the port DFTs.
//
// a = 0.5*(V + Z*I)/sqrt(Z), b = 0.5*(V - Z*I)/sqrt(Z) for a real port impedance Z,
// then S(out, in) = b(out)/a(in) for every source port `in`, with the reference's
// guard (|a|^2 < 1e-300 gives S = 0), 20*log10|S| and atan2 in degrees
// (js_fdtd/src/sparameters.js:23-101).  Each expression keeps the reference's
// operation order so results agree bit for bit under the same Math library.

export function computePortWaves(ports, sampledVoltages, sampledCurrents) {
  for (const port of ports) {
    const obsV = sampledVoltages[port.sampled_voltage_index - 1];
    const obsI = sampledCurrents[port.sampled_current_index - 1];
    const Z = typeof port.impedance === 'number' ? port.impedance : port.impedance.re;
    const root = Math.sqrt(Z);
    const V = obsV.frequency_domain_value, I = obsI.frequency_domain_value;
    const M = V.re.length;
    const a = { re: new Float64Array(M), im: new Float64Array(M) };
    const b = { re: new Float64Array(M), im: new Float64Array(M) };
    for (let m = 0; m < M; m++) {
      a.re[m] = 0.5 * (V.re[m] + Z * I.re[m]) / root;
      a.im[m] = 0.5 * (V.im[m] + Z * I.im[m]) / root;
      b.re[m] = 0.5 * (V.re[m] - Z * I.re[m]) / root;
      b.im[m] = 0.5 * (V.im[m] - Z * I.im[m]) / root;
    }
    port.a = a;
    port.b = b;
    port.frequencies = obsV.frequencies;
  }
}

function ratio(b, a) {
  const M = a.re.length;
  const out = { re: new Float64Array(M), im: new Float64Array(M),
                mag_dB: new Float64Array(M), phase_deg: new Float64Array(M) };
  for (let m = 0; m < M; m++) {
    const ar = a.re[m], ai = a.im[m], br = b.re[m], bi = b.im[m];
    const den = ar * ar + ai * ai;
    const sr = den < 1e-300 ? 0 : (br * ar + bi * ai) / den;
    const si = den < 1e-300 ? 0 : (bi * ar - br * ai) / den;
    const mag = Math.sqrt(sr * sr + si * si);
    out.re[m] = sr;
    out.im[m] = si;
    out.mag_dB[m] = mag > 0 ? 20 * Math.log10(mag) : -Infinity;
    out.phase_deg[m] = Math.atan2(si, sr) * 180 / Math.PI;
  }
  return out;
}

export function computeSMatrix(ports) {
  for (const src of ports) {
    if (!src.is_source_port) continue;
    src.S = ports.map((out) => ratio(out.b, src.a));
  }
}

/** Port waves then S-matrix (sparameters.js:110-116). */
export function postProcessSParameters({ ports, sampledVoltages, sampledCurrents }) {
  computePortWaves(ports, sampledVoltages, sampledCurrents);
  computeSMatrix(ports);
}
