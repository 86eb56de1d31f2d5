"""Regenerate tests/golden/setup_*.npz from the reference's host-side setup code.

Runs `make_setup_golden.js` under node on the reference source directory (in the
development container only; the reference is never copied) and packs the
numeric outputs into compressed .npz fixtures plus one JSON index.

usage: python tests/golden/make_setup_golden.py [/path/to/reference/source]
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/source"
    with tempfile.TemporaryDirectory() as raw:
        subprocess.run(["node", os.path.join(HERE, "make_setup_golden.js"), src, raw], check=True)
        with open(os.path.join(raw, "index.json")) as f:
            index = json.load(f)
        kinds = {"u8": np.uint8, "f32": np.float32, "f64": np.float64}
        for name in [k for k in index if k not in ("atmosphere", "theory")]:
            arrays = {}
            for fn in sorted(os.listdir(raw)):
                if not fn.startswith(name + "."):
                    continue
                _, field, kind = fn.split(".")
                arrays[field] = np.fromfile(os.path.join(raw, fn), dtype=kinds[kind])
            np.savez_compressed(os.path.join(HERE, f"setup_{name}.npz"), **arrays)
        with open(os.path.join(HERE, "setup_index.json"), "w") as f:
            json.dump(index, f, indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
