"""Write tests/golden/fixtures/published-ifa-fr4-substrate-200.json from the only run of
the reference's own engine that ships with it: js_fdtd/results/out.json.

That file is the output of `node src/index.js` (js_fdtd/src/index.js:108-130 schema) for
the scenario js_fdtd/scenarios/ifa-fr4-substrate.json cut to 200 steps (its meta block
says so), computed by the multithreaded FP64 WASM engine on the author's machine (Node
v22.15.0, 15 workers; js_fdtd/out.txt:58, js_fdtd/out2.txt:11-25).  The fixture keeps
the data only -- the sampled voltage trace, S11 magnitude and phase, and the four
far-field cuts -- as the exact doubles the JSON holds, plus the meta block.

    python tests/golden/gen_published.py [/path/to/reference]
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "js_fdtd", "results", "out.json")
    with open(src) as f:
        d = json.load(f)
    meta = {k: d["meta"][k] for k in ("generated", "scenario", "numberOfTimeSteps", "dt")}
    out = {
        "provenance": "js_fdtd/results/out.json of the reference (scenario ifa-fr4-substrate, 200 steps, "
                      "FP64 wasm-cpu engine with 15 workers); data only",
        "meta": meta,
        "frequencies_Hz": d["meta"]["frequencies_Hz"],
        "v": d["timeDomain"]["voltages"][0]["values"],
        "v_label": d["timeDomain"]["voltages"][0]["label"],
        "s11_dB": d["sparams"][0]["magnitude_dB"],
        "s11_phase_deg": d["sparams"][0]["phase_deg"],
        "farfield": [{"plane": c["plane"], "freq_GHz": c["freq_GHz"], "dataTheta_dB": c["dataTheta_dB"],
                      "dataPhi_dB": c["dataPhi_dB"]} for c in d["farfield"]],
    }
    dst = os.path.join(HERE, "fixtures", "published-ifa-fr4-substrate-200.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {dst}")


if __name__ == "__main__":
    main()
