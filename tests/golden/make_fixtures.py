"""Writes tests/golden/reference_results.json: the hot-path entries of the
reference's own recorded test results (tests/results-cpu.json, written by
tests/run-tests.mjs --json; tests/results-gpu.json, written by
tests/run-gpu-browser.mjs). Data only: recorded inputs' outputs, no source.

Run in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import sys

CPU_KEYS = ["stoker", "ritter", "lakeSubmerged", "lakeIsland", "thacker", "rainMass", "urban", "valleyDam"]
GPU_KEYS = ["stoker", "ritter", "lakeSubmerged", "lakeIsland", "thacker", "thackerMass", "rainMass", "urban",
            "valleyDam", "inlets", "tiles", "manyTiles", "rerun", "lts", "hazard", "stage", "barriers", "gi", "greenAmptScour"]


def main(ref):
    cpu = json.load(open(os.path.join(ref, "tests", "results-cpu.json")))
    gpu = json.load(open(os.path.join(ref, "tests", "results-gpu.json")))
    out = {"cpu": {k: cpu[k] for k in CPU_KEYS}, "gpu": {k: gpu[k] for k in GPU_KEYS},
           "source": {"cpu": "tests/results-cpu.json", "gpu": "tests/results-gpu.json"}}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_results.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
