"""Writes tests/golden/fixtures.json: oracle outputs of the hand-built fixture
scenes (coverage counts and SHA-256 of the visibility buffer and shadow maps).

The reference has no tests or fixtures of its own; these values are produced
by this project's oracle and serve as regression anchors for the oracle and
as bit-exact targets for the HIP path.  Run from the repository root:
    python -m tests.golden.make_golden
"""
import hashlib
import json
from pathlib import Path

import numpy as np

from lighthugger_amd import abi
from lighthugger_amd.scene import make_uniforms
from oracle.compare import oracle_frame
from tests.fixtures import fixtures

OUT = Path(__file__).resolve().parent / "fixtures.json"


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def summarize(snap: dict) -> dict:
    vis = snap["visbuffer"]
    return {
        "covered_pixels": int((vis != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum()),
        "visbuffer_sha256": digest(vis),
        "shadow_sha256": digest(snap["shadow"]),
        "draws": [int(x) for x in snap["counters"][:5]],
        "min_depth": int(snap["misc"]["min_depth"]),
        "max_depth": int(snap["misc"]["max_depth"]),
        "radiance_mean": [float(x) for x in snap["radiance"][..., :3].mean(axis=(0, 1))],
    }


def main():
    res = {}
    for name, (scene, cam, cfg) in fixtures().items():
        res[name] = summarize(oracle_frame(scene, cfg, make_uniforms(cam, scene.num_instances)))
        print(name, res[name]["covered_pixels"], res[name]["draws"])
    OUT.write_text(json.dumps(res, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
