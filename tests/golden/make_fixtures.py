"""Writes the reference fixtures of tests/golden from the REFERENCE's own
kernels, built from its source by `make -C oracle ref` (this container only).

  ref_small.npz    m=n=p=3, buff_size=6, v_size=2 (a 16^3 `is` box): the input
                   field (seeded recipe 2: sine modes, a step, 1e3 jump,
                   exactly-constant boxes) and the reference's vL/vR for the
                   x, y and z sequences of weno3 / weno5 / weno7 / mp5 / mp7
                   (order-1 copy, then the kernel, then MP for mp*).
                   weno3 halo cells are stored but are undefined in the
                   reference (out-of-bounds coefficient reads).
  ref_branches.npz the reference's TENO, WENO-Z and mapped-WENO weights: the
                   reference built with its compile-time parameter teno,
                   wenoz or mapped_weno set (oracle/_ref/<branch>/) and run
                   with wenojs=F, m=n=p=3, v_size=1 (seeded recipe 2), weno3
                   (not TENO: the reference has no TENO3 branch) / weno5 /
                   weno7, core range of vL/vR for x, y and z.
  ref_checksums.json  X/Y/Z checksums (sum|vL|+sum|vR|) of the reference at
                   m=n=p=31, v_size=8, for the harness field and for a seeded
                   field, per kernel sequence.

Run: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from wenostandalone_amd import fields  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    if not oracle.have_ref():
        sys.exit("build the reference first: make -C oracle ref")
    m, v, b = 3, 2, 6
    f = fields.seeded_field(m, m, m, v, b, seed=11, recipe=2)
    out = {"field": f, "m": m, "v": v, "buff": b}
    for mode in ["weno3", "weno5", "weno7", "mp5", "mp7"]:
        r = oracle.run_ref(mode, m, v, b, f)
        for d, name in enumerate("xyz"):
            out[f"{mode}_{name}_L"] = r[2 * d]
            out[f"{mode}_{name}_R"] = r[2 * d + 1]
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), **out)

    m, v = 3, 1
    f = fields.seeded_field(m, m, m, v, b, seed=13, recipe=2)
    out = {"field": f, "m": m, "v": v, "buff": b}
    for branch in oracle.BRANCHES:
        for mode in (["weno5", "weno7"] if branch == "teno" else ["weno3", "weno5", "weno7"]):
            r = oracle.run_ref(mode, m, v, b, f, wenojs=False, branch=branch)
            for d, name in enumerate("xyz"):
                out[f"{branch}_{mode}_{name}_L"] = np.ascontiguousarray(r[2 * d][3:-3])
                out[f"{branch}_{mode}_{name}_R"] = np.ascontiguousarray(r[2 * d + 1][3:-3])
    np.savez_compressed(os.path.join(HERE, "ref_branches.npz"), **out)

    sums = {}
    m, v = 31, 8
    for fname, field in [("harness", None), ("seeded0", fields.seeded_field(m, m, m, v, b, seed=5))]:
        for mode in ["copy", "weno5", "weno7", "mp5", "mp7"]:
            r = oracle.run_ref(mode, m, v, b, field)
            sums[f"{fname}_{mode}"] = [float(np.abs(r[2 * d]).sum() + np.abs(r[2 * d + 1]).sum())
                                       for d in range(3)]
    with open(os.path.join(HERE, "ref_checksums.json"), "w") as fh:
        json.dump({"m": m, "v": v, "buff": b, "seeded_seed": 5, "checksums_xyz": sums}, fh, indent=1)


if __name__ == "__main__":
    main()
