"""Write the seeded synthetic scenarios (js_fdtd scenario schema) used for measurement.

    python special-topics_amd/scenarios/gen_synthetic.py

roofline-256: 256^3 cells of 1 mm including an 8-cell CPML on every face; 20 random
dielectric boxes (eps_r ~ U[1,10], sigma_e ~ U[0,0.01] S/m, edges ~ U[8,64] cells,
placed uniformly inside the 240^3 interior) drawn from MT19937(seed=1); a 1-cell
z-directed voltage source (Hertzian dipole) at the centre with a Gaussian pulse, a
port voltage/current observer on it; 2000 steps.

patch-512: 512x512x384 cells, 10-cell CPML, 5 random substrate layers (eps_r ~
U[2,12], 4-20 cells thick) on a PEC ground plane, a 60x40-cell PEC patch fed by a
lumped port, NTFF box 5 cells inside the CPML at 16 frequencies, 8000 steps.
"""

from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def cpml_boundary(nc: int, air=0) -> dict:
    b = {}
    for f in ("xn", "xp", "yn", "yp", "zn", "zp"):
        b["type_" + f] = "cpml"
    for f in ("xn", "xp", "yn", "yp", "zn", "zp"):
        b["air_buffer_" + f] = air
    for f in ("xn", "xp", "yn", "yp", "zn", "zp"):
        b["cpml_cells_" + f] = nc
    b.update({"cpml_order": 3, "cpml_sigma_factor": 1.3, "cpml_kappa_max": 7,
              "cpml_alpha_min": 0, "cpml_alpha_max": 0.05})
    return b


def roofline_256() -> dict:
    rng = np.random.Generator(np.random.MT19937(1))
    d = 1e-3
    inner = 240
    mats = [{"eps_r": 1, "mu_r": 1, "sigma_e": 0, "sigma_m": 0}]
    bricks = [{"min_x": 0, "min_y": 0, "min_z": 0, "max_x": inner * d, "max_y": inner * d,
               "max_z": inner * d, "material_type": 1}]
    for _ in range(20):
        edge = rng.integers(8, 65, size=3)
        lo = [int(rng.integers(0, inner - e + 1)) for e in edge]
        mats.append({"eps_r": float(rng.uniform(1, 10)), "mu_r": 1,
                     "sigma_e": float(rng.uniform(0, 0.01)), "sigma_m": 0})
        bricks.append({"min_x": lo[0] * d, "min_y": lo[1] * d, "min_z": lo[2] * d,
                       "max_x": (lo[0] + int(edge[0])) * d, "max_y": (lo[1] + int(edge[1])) * d,
                       "max_z": (lo[2] + int(edge[2])) * d, "material_type": len(mats)})
    c = inner // 2 * d
    return {
        "name": "roofline-256",
        "description": "256^3 FP32 stencil roofline case: free space with 20 seeded random dielectric "
                       "boxes, 8-cell CPML, centre Hertzian dipole (1-cell z voltage source), 2000 steps.",
        "numberOfTimeSteps": 2000, "courantFactor": 0.9, "numberOfCellsPerWavelength": 20,
        "cell": {"dx": d, "dy": d, "dz": d},
        "boundary": cpml_boundary(8),
        "materialTypes": mats, "bricks": bricks, "spheres": [],
        "waveforms": {"gaussian": [{"number_of_cells_per_wavelength": 0}]},
        "voltageSources": [{"min_x": c, "min_y": c, "min_z": c, "max_x": c, "max_y": c, "max_z": c + d,
                            "direction": "zp", "resistance": 50, "magnitude": 1,
                            "waveform_type": "gaussian", "waveform_index": 1}],
        "sampledVoltages": [{"min_x": c, "min_y": c, "min_z": c, "max_x": c, "max_y": c, "max_z": c + d,
                             "direction": "zp", "label": "v1"}],
        "sampledCurrents": [{"min_x": c - d, "min_y": c - d, "min_z": c + d / 2, "max_x": c + d,
                             "max_y": c + d, "max_z": c + d / 2, "direction": "zp", "label": "i1"}],
        "ports": [{"sampled_voltage_index": 1, "sampled_current_index": 1, "impedance": 50,
                   "is_source_port": True}],
        "farfield": {"frequencies": [], "number_of_cells_from_outer_boundary": 13},
        "dft": {"maxFreq": 10e9, "step": 20e6},
    }


def patch_512() -> dict:
    rng = np.random.Generator(np.random.MT19937(3))
    d = 1e-3
    nx_in, ny_in, nz_in = 492, 492, 364          # interior inside a 10-cell CPML
    mats = [{"eps_r": 1, "mu_r": 1, "sigma_e": 0, "sigma_m": 0},
            {"eps_r": 1, "mu_r": 1, "sigma_e": 1e10, "sigma_m": 0}]
    bricks = [{"min_x": 0, "min_y": 0, "min_z": 0, "max_x": nx_in * d, "max_y": ny_in * d,
               "max_z": nz_in * d, "material_type": 1}]
    z = 40
    bricks.append({"min_x": 0, "min_y": 0, "min_z": z * d, "max_x": nx_in * d, "max_y": ny_in * d,
                   "max_z": z * d, "material_type": 2})                       # ground plane
    for _ in range(5):
        t = int(rng.integers(4, 21))
        mats.append({"eps_r": float(rng.uniform(2, 12)), "mu_r": 1, "sigma_e": 0, "sigma_m": 0})
        bricks.append({"min_x": 0, "min_y": 0, "min_z": z * d, "max_x": nx_in * d, "max_y": ny_in * d,
                       "max_z": (z + t) * d, "material_type": len(mats)})
        z += t
    px0, py0 = nx_in // 2 - 30, ny_in // 2 - 20
    bricks.append({"min_x": px0 * d, "min_y": py0 * d, "min_z": z * d, "max_x": (px0 + 60) * d,
                   "max_y": (py0 + 40) * d, "max_z": z * d, "material_type": 2})  # patch
    fx, fy = px0 + 15, py0 + 20
    return {
        "name": "patch-512",
        "description": "512x512x384 FP32 patch antenna on a seeded 5-layer substrate, 10-cell CPML, "
                       "lumped port, NTFF at 16 frequencies, 8000 steps.",
        "numberOfTimeSteps": 8000, "courantFactor": 0.9, "numberOfCellsPerWavelength": 20,
        "cell": {"dx": d, "dy": d, "dz": d},
        "boundary": cpml_boundary(10),
        "materialTypes": mats, "bricks": bricks, "spheres": [],
        "waveforms": {"derivative_gaussian": [{"number_of_cells_per_wavelength": 30}]},
        "voltageSources": [{"min_x": fx * d, "min_y": fy * d, "min_z": 40 * d, "max_x": fx * d,
                            "max_y": fy * d, "max_z": z * d, "direction": "zp", "resistance": 50,
                            "magnitude": 1, "waveform_type": "derivative_gaussian", "waveform_index": 1}],
        "sampledVoltages": [{"min_x": fx * d, "min_y": fy * d, "min_z": 40 * d, "max_x": fx * d,
                             "max_y": fy * d, "max_z": z * d, "direction": "zp", "label": "v1"}],
        "sampledCurrents": [{"min_x": (fx - 1) * d, "min_y": (fy - 1) * d, "min_z": (z - 1) * d,
                             "max_x": (fx + 1) * d, "max_y": (fy + 1) * d, "max_z": (z - 1) * d,
                             "direction": "zp", "label": "i1"}],
        "ports": [{"sampled_voltage_index": 1, "sampled_current_index": 1, "impedance": 50,
                   "is_source_port": True}],
        "farfield": {"frequencies": [float(f) for f in np.linspace(1e9, 8.5e9, 16)],
                     "number_of_cells_from_outer_boundary": 15},
        "dft": {"maxFreq": 10e9, "step": 10e6},
    }


def main():
    for sc in (roofline_256(), patch_512()):
        path = os.path.join(HERE, sc["name"] + ".json")
        with open(path, "w") as f:
            json.dump(sc, f, indent=1)
        print("wrote", path)


if __name__ == "__main__":
    main()
