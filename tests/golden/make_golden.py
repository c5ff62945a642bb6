"""Writes tests/golden/golden.json: regression vectors for the RNG, seeding, multiplier table and a
tiny transport run, produced by the CPU oracle (C++).  The reference ships no golden vectors or
fixtures for this path, so these pin the oracle against later changes; what pins the oracle
against the reference's semantics is the closed-form, property and known-answer tests.

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
import ymc3d_amd as y  # noqa: E402


def tiny_scene():
    """8x8x6 voxels of 0.02: two layers (n 1.4 / 1.33) with a 2x2x2 absorbing block, ambient below."""
    vals = np.zeros((6, 8, 8), dtype=np.int32)
    vals[2:5] = 1
    vals[2:4, 3:5, 3:5] = 2
    vals[5] = 255
    rec = np.array([[[2.0, 80.0, 0.85, 1.40], [0.5, 40.0, 0.80, 1.33], [30.0, 60.0, 0.90, 1.40]]])
    return vals, rec


def main():
    table = y.generate_multipliers(65)
    out = {"multipliers_first_65": [int(v) for v in table]}

    x0 = np.array([1, 0x0000000100000000, 0x12345678_9ABCDEF0 & 0x00FFFFFF_FFFFFFFF, (int(table[4]) - 2) << 32 | 0xFFFFFFFE],
                  dtype=np.uint64)
    a = table[1:5].copy()
    co, xf = O.mwc_fill(x0, a, 16, 0)
    oc, _ = O.mwc_fill(x0, a, 16, 1)
    out["mwc"] = {"x0": [int(v) for v in x0], "a": [int(v) for v in a],
                  "co_bits": co.view(np.uint32).T.tolist(), "oc_bits": oc.view(np.uint32).T.tolist(),
                  "x_final": [int(v) for v in xf]}

    st, xs, as_ = O.init_rng(table, 20241203, 100)
    assert st == 0
    out["seeding"] = {"seed": 20241203, "table_len": 65, "x": [int(v) for v in xs], "a": [int(v) for v in as_]}

    vals, rec = tiny_scene()
    rows = np.where(vals == 255, 0, vals + 1).astype(np.uint8)
    L, T, _ = rec.shape
    props = []
    for q, amb in enumerate((0.0, 0.0, 0.0, 1.0)):
        arr = np.empty((L, T + 1), dtype=np.float32)
        arr[:, 0] = amb
        arr[:, 1:] = rec[:, :, q]
        props.append(arr)
    chains = O.build_chains([3000], 64)
    st, xc, ac = O.init_rng(table, 99, chains.size)
    res = O.run(rows, (0.02, 0.02, 0.02), props, chains, xc, ac, illumination=1, num_photons=3000)
    out["tiny_run"] = {
        "file_values": vals.tolist(), "records": rec.tolist(), "voxel": 0.02, "photons": 3000, "seed": 99,
        "chain_lanes": 64, "illumination": "pencil",
        "R_w": res.R_w[0].tolist(), "R_n": res.R_n[0].tolist(),
        "totals": {k: int(res.totals[0][k]) for k in res.totals.dtype.names},
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
