// Scenario JSON -> simulation problem.
//
// Same input schema, defaults and set-up order as js_fdtd/src/scenario.js:32-131:
// grid, materials, PEC plates, general coefficients, CPML (kappa division), waveforms,
// voltage sources, lumped-element coefficient overrides, observers, NTFF box.  The
// order matters: the lumped overrides come after the kappa division, CPsi before it.
// Extension: `resistors` and `capacitors` lists are honoured (coefficient overrides
// only, coefficients.js:243-325); the reference loader ignores them.

import fs from 'fs';
import path from 'path';
import { sizeDomain, allocateFields } from './grid.mjs';
import { buildMaterialGrid, computeMaterialComponents, applyPECPlates } from './materials.mjs';
import { computeGeneralCoefficients, applyLumpedElementCoefficients, initPassiveElement } from './coefficients.mjs';
import { initCPML } from './cpml.mjs';
import { initWaveforms, initVoltageSources } from './sources.mjs';
import {
  initSampledVoltages, initSampledCurrents, initSampledElectricFields, initSampledMagneticFields,
} from './samplers.mjs';
import { initFarfield } from './farfield.mjs';

export const DEFAULTS = {
  courantFactor: 0.9,
  numberOfCellsPerWavelength: 20,
  numberOfTimeSteps: 4000,
  cell: { dx: 0.262e-3, dy: 0.4e-3, dz: 0.4e-3 },
  dft: { maxFreq: 10e9, step: 20e6 },
  farfield: { frequencies: [], number_of_cells_from_outer_boundary: 13 },
};

const HERE = path.dirname(new URL(import.meta.url).pathname);
export const SCENARIOS_DIR = path.resolve(HERE, '../../scenarios');

export function listScenarios() {
  if (!fs.existsSync(SCENARIOS_DIR)) return [];
  return fs.readdirSync(SCENARIOS_DIR).filter((f) => f.endsWith('.json'))
    .map((f) => f.replace(/\.json$/, '')).sort();
}

/** Parse a scenario by name (looked up in SCENARIOS_DIR) or by path to a .json file. */
export function loadScenario(nameOrPath) {
  let file = nameOrPath;
  if (!file.endsWith('.json')) file = path.join(SCENARIOS_DIR, nameOrPath + '.json');
  else if (!path.isAbsolute(file)) file = path.resolve(file);
  if (!fs.existsSync(file)) {
    throw new Error(`Scenario not found: ${nameOrPath}. Available: ${listScenarios().join(', ') || '(none)'}`);
  }
  return JSON.parse(fs.readFileSync(file, 'utf8'));
}

function pick(v, dflt) { return v === undefined || v === null ? dflt : v; }

/**
 * Build the full FP64 problem.  options.steps overrides the step count (as does the
 * FDTD_STEPS environment variable, which takes precedence like scenario.js:76-78).
 */
export function buildProblemFromScenario(scenario, options) {
  const opts = options || {};
  const courantFactor = pick(scenario.courantFactor, DEFAULTS.courantFactor);
  const cellsPerWavelength = pick(scenario.numberOfCellsPerWavelength, DEFAULTS.numberOfCellsPerWavelength);
  const numberOfTimeSteps = Number(process.env.FDTD_STEPS) || opts.steps ||
    scenario.numberOfTimeSteps || DEFAULTS.numberOfTimeSteps;
  const cell = Object.assign({}, DEFAULTS.cell, scenario.cell || {});
  const { dx, dy, dz } = cell;
  const boundary = scenario.boundary;
  const materialTypes = scenario.materialTypes;
  const bricks = scenario.bricks || [];
  const spheres = scenario.spheres || [];

  const grid = sizeDomain({ dx, dy, dz, boundary, bricks, spheres, courantFactor, numberOfTimeSteps });
  grid.fields = allocateFields(grid);

  const matGrid = buildMaterialGrid(grid, bricks, spheres);
  const matComps = computeMaterialComponents(matGrid, materialTypes, grid);
  applyPECPlates(bricks, materialTypes, matComps, grid);

  const coeffs = computeGeneralCoefficients(matComps, grid);
  const cpml = initCPML(boundary, coeffs, grid);

  const waveforms = initWaveforms(scenario.waveforms || {}, numberOfTimeSteps, grid.dt,
                                  cellsPerWavelength, [dx, dy, dz]);
  const voltageSources = scenario.voltageSources || [];
  initVoltageSources(voltageSources, waveforms, grid);
  const resistors = scenario.resistors || [];
  const capacitors = scenario.capacitors || [];
  for (const r of resistors) initPassiveElement(r, grid, 'resistor');
  for (const c of capacitors) initPassiveElement(c, grid, 'capacitor');
  applyLumpedElementCoefficients(coeffs, matComps, grid, voltageSources, resistors, capacitors);

  const sampledVoltages = scenario.sampledVoltages || [];
  const sampledCurrents = scenario.sampledCurrents || [];
  const sampledEFields = scenario.sampledElectricFields || [];
  const sampledHFields = scenario.sampledMagneticFields || [];
  initSampledVoltages(sampledVoltages, grid);
  initSampledCurrents(sampledCurrents, grid);
  initSampledElectricFields(sampledEFields, grid);
  initSampledMagneticFields(sampledHFields, grid);

  const ffParams = Object.assign({}, DEFAULTS.farfield, scenario.farfield || {});
  const farfield = initFarfield(ffParams, grid);
  const dft = Object.assign({}, DEFAULTS.dft, scenario.dft || {});

  return {
    grid, coeffs, cpml,
    samplers: { sampledVoltages, sampledCurrents, sampledEFields, sampledHFields },
    sources: { voltageSources, currentSources: [], inductors: [], diodes: [] },
    farfield, ports: scenario.ports || [], frequencies: ffParams.frequencies || [],
    numberOfTimeSteps, dt: grid.dt, dft,
    meta: { name: scenario.name, description: scenario.description },
    materialGrid: matGrid,
  };
}
