// Port-quantity DFT after the time loop, FP64.
//
// X(f) = dt * sum_n x_n exp(-j w ((n+1) dt + shift)), as js_fdtd/src/dft.js:17-38
// evaluates it (sequential sums, cos/sin of the full phase); shift = 0 for voltages
// and -dt/2 for currents (js_fdtd/src/index.js:69-70).  The frequency grid is
// f_i = (i+1)*step, i < floor(maxFreq/step) (index.js:64-67).

export function frequencyGrid(dft) {
  const maxFreq = dft && dft.maxFreq != null ? dft.maxFreq : 10e9;
  const step = dft && dft.step != null ? dft.step : 20e6;
  const n = Math.floor(maxFreq / step);
  const f = new Float64Array(n);
  for (let i = 0; i < n; i++) f[i] = (i + 1) * step;
  return f;
}

export function timeToFreqDomain(x, dt, freqArr, timeShift) {
  const N = x.length, M = freqArr.length;
  const Xre = new Float64Array(M), Xim = new Float64Array(M);
  for (let m = 0; m < M; m++) {
    const w = 2 * Math.PI * freqArr[m];
    let re = 0, im = 0;
    for (let n = 0; n < N; n++) {
      const phase = w * ((n + 1) * dt + timeShift);
      re += x[n] * Math.cos(phase);
      im -= x[n] * Math.sin(phase);
    }
    Xre[m] = re * dt;
    Xim[m] = im * dt;
  }
  return { re: Xre, im: Xim };
}

/** Attach frequency_domain_value / frequencies to each observer (dft.js:90-95). */
export function computeObserverDFT(observers, dt, timeShift, freqArr) {
  for (const obs of observers) {
    obs.frequency_domain_value = timeToFreqDomain(obs.sampled_value, dt, freqArr, timeShift);
    obs.frequencies = freqArr;
  }
}
