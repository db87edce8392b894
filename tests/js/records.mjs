// Shared by the test drivers: the post-processing record (dft, s11, farfield) of a
// finished run, in the layout of the golden run records (base64 little-endian f64).

import crypto from 'crypto';

import { frequencyGrid, computeObserverDFT } from '../../special-topics_amd/js/post/dft.mjs';
import { postProcessSParameters } from '../../special-topics_amd/js/post/sparams.mjs';
import { calcRadiatedPower, calcDirectivity, standardCuts } from '../../special-topics_amd/js/post/farfield.mjs';

export const b64 = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64');

export const f64 = (s) => {
  const b = Buffer.from(s, 'base64');
  return new Float64Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
};

export const sha = (a) => crypto.createHash('sha256').update(Buffer.from(a.buffer, a.byteOffset, a.byteLength)).digest('hex');

/** Port DFTs and S11 of the samplers' traces, plus the far-field cuts when withFarfield. */
export function postRecord(sim, withFarfield) {
  const out = {};
  const { sampledVoltages: sv, sampledCurrents: si } = sim.samplers;
  const freqArr = frequencyGrid(sim.dft);
  const dt = sim.grid.dt;
  computeObserverDFT(sv, dt, 0, freqArr);
  computeObserverDFT(si, dt, -dt / 2, freqArr);
  out.dft = { v: sv.map((o) => [b64(o.frequency_domain_value.re), b64(o.frequency_domain_value.im)]),
              i: si.map((o) => [b64(o.frequency_domain_value.re), b64(o.frequency_domain_value.im)]) };
  if (sim.ports.length) {
    postProcessSParameters({ ports: sim.ports, sampledVoltages: sv, sampledCurrents: si, dt, freqArr });
    out.s11 = sim.ports.filter((p) => p.is_source_port).map((p) => ({
      re: b64(p.S[0].re), im: b64(p.S[0].im), dB: b64(p.S[0].mag_dB), phase: b64(p.S[0].phase_deg) }));
  }
  if (withFarfield && sim.farfield.nFreq > 0) {
    const ff = sim.farfield;
    const t0 = Date.now();
    const prad = calcRadiatedPower(ff, sim.grid);
    const cuts = standardCuts();
    const xy = calcDirectivity(ff, sim.grid, prad, cuts.xy.theta, cuts.xy.phi);
    const xz = calcDirectivity(ff, sim.grid, prad, cuts.xz.theta, cuts.xz.phi);
    out.farfield = { prad: b64(prad),
                     xy: xy.dataTheta.map((a, f) => [b64(a), b64(xy.dataPhi[f])]),
                     xz: xz.dataTheta.map((a, f) => [b64(a), b64(xz.dataPhi[f])]) };
    out.farfield_s = (Date.now() - t0) / 1000;
  }
  return out;
}
