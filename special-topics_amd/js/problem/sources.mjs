// Excitation waveforms and voltage-source set-up.
//
// Waveform definitions follow js_fdtd/src/sources.js:16-94 (samples at
// t_n = (n + 0.5) dt; Gaussian, derivative-of-Gaussian, cosine-modulated Gaussian,
// sinusoid and unit step), voltage sources sources.js:104-145 (the per-edge voltage
// v_n = sign * magnitude / n_edges * waveform_n and the per-edge resistance).

import { C0 } from './constants.mjs';

function sampleTimes(n, dt) {
  const t = new Float64Array(n);
  for (let i = 0; i < n; i++) t[i] = (i + 0.5) * dt;
  return t;
}

function pulseWidth(w, defaultCells, maxCell) {
  const nc = w.number_of_cells_per_wavelength === 0 ? defaultCells : w.number_of_cells_per_wavelength;
  w.maximum_frequency = C0 / (nc * maxCell);
  const tau = (nc * maxCell) / (2 * C0);
  w.tau = tau;
  w.t_0 = 4.5 * tau;
  return tau;
}

/** Attach `.waveform` (Float64Array of numberOfTimeSteps samples) to every entry. */
export function initWaveforms(waveforms, numberOfTimeSteps, dt, defaultCells, cellSizes) {
  const maxCell = Math.max(...cellSizes);
  const time = sampleTimes(numberOfTimeSteps, dt);
  const N = numberOfTimeSteps;

  for (const w of waveforms.sinusoidal || []) {
    w.waveform = new Float64Array(N);
    for (let n = 0; n < N; n++) w.waveform[n] = Math.sin(2 * Math.PI * w.frequency * time[n]);
    w.t_0 = 0;
  }
  for (const w of waveforms.unit_step || []) {
    w.waveform = new Float64Array(N).fill(1);
    for (let n = 0; n < w.start_time_step - 1; n++) w.waveform[n] = 0;
    w.t_0 = 0;
  }
  for (const w of waveforms.gaussian || []) {
    const tau = pulseWidth(w, defaultCells, maxCell);
    w.waveform = new Float64Array(N);
    for (let n = 0; n < N; n++) {
      const arg = (time[n] - w.t_0) / tau;
      w.waveform[n] = Math.exp(-arg * arg);
    }
  }
  for (const w of waveforms.derivative_gaussian || []) {
    const tau = pulseWidth(w, defaultCells, maxCell);
    const scale = -Math.sqrt(2 * Math.E) / tau;
    w.waveform = new Float64Array(N);
    for (let n = 0; n < N; n++) {
      const s = time[n] - w.t_0;
      const arg = s / tau;
      w.waveform[n] = scale * s * Math.exp(-arg * arg);
    }
  }
  for (const w of waveforms.cosine_modulated_gaussian || []) {
    const tau = 0.966 / w.bandwidth;
    w.tau = tau;
    w.t_0 = 4.5 * tau;
    w.waveform = new Float64Array(N);
    for (let n = 0; n < N; n++) {
      const s = time[n] - w.t_0;
      const arg = s / tau;
      w.waveform[n] = Math.cos(2 * Math.PI * w.modulation_frequency * s) * Math.exp(-arg * arg);
    }
  }
  return waveforms;
}

/** 1-based box, per-edge resistance and per-edge voltage of each voltage source. */
export function initVoltageSources(voltageSources, waveforms, grid) {
  const { dx, dy, dz, min_x, min_y, min_z } = grid;
  for (const vs of voltageSources) {
    vs.is = Math.round((vs.min_x - min_x) / dx) + 1;
    vs.js = Math.round((vs.min_y - min_y) / dy) + 1;
    vs.ks = Math.round((vs.min_z - min_z) / dz) + 1;
    vs.ie = Math.round((vs.max_x - min_x) / dx) + 1;
    vs.je = Math.round((vs.max_y - min_y) / dy) + 1;
    vs.ke = Math.round((vs.max_z - min_z) / dz) + 1;

    const ext = [1 + vs.ie - vs.is, 1 + vs.je - vs.js, 1 + vs.ke - vs.ks];
    const gap = [vs.ie - vs.is, vs.je - vs.js, vs.ke - vs.ks];
    const axis = 'xyz'.indexOf(vs.direction[0]);
    const o1 = axis === 0 ? 1 : 0, o2 = axis === 2 ? 1 : 2;
    const nFields = gap[axis];
    const rFactor = ext[o1] * ext[o2] / gap[axis];

    const sign = vs.direction[1] === 'n' ? -1 : 1;
    const vFactor = sign * vs.magnitude / nFields;
    vs.resistance_per_component = rFactor * vs.resistance;

    const base = waveforms[vs.waveform_type][vs.waveform_index - 1].waveform;
    vs.voltage_per_e_field = new Float64Array(base.length);
    vs.waveform = new Float64Array(base.length);
    for (let n = 0; n < base.length; n++) {
      vs.voltage_per_e_field[n] = vFactor * base[n];
      vs.waveform[n] = vFactor * nFields * base[n];
    }
  }
}
