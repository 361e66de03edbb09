"""Writes tests/golden/hsfc_keys.json: Hilbert keys computed by the
reference's own fhsfc2d_ (src/main/hilbert.C, built from the reference
sources by `make -C oracle ref` into oracle/_ref/).  Run in a container that
has the reference; the JSON (inputs and expected outputs only) is committed.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import flatten_oracle  # noqa: E402


def main():
    rng = np.random.default_rng(20240611)
    special = [0.0, 1.0, 0.5, 0.25, 0.75, -0.5, 1.5, 1e-12, 1 - 1e-12, 2.0 ** -32, 0.1, 1 / 3]
    pts = [(a, b) for a in special for b in special]
    pts += [tuple(p) for p in rng.random((200, 2))]
    # dyadic centres of a 64 x 48 grid (the flattening's typical inputs)
    for _ in range(100):
        i, j = rng.integers(0, 64), rng.integers(0, 48)
        pts.append(((i + 0.5) / 64, (j + 0.5) / 48))
    xy = np.array(pts, dtype=np.float64)
    keys = flatten_oracle.ref_keys(xy)
    if keys is None:
        raise SystemExit("oracle/_ref/libhsfc_ref.so missing: run `make -C oracle ref` first")
    out = {"source": "fhsfc2d_ of src/main/hilbert.C (nkey = 2)",
           "norm_xy": [[float(a), float(b)] for a, b in xy],
           "keys": [[int(a), int(b)] for a, b in keys]}
    with open(os.path.join(HERE, "hsfc_keys.json"), "w") as f:
        json.dump(out, f)
    print(f"wrote {len(xy)} key vectors")


if __name__ == "__main__":
    main()
