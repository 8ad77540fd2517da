#!/usr/bin/env python3
"""Write the whole-step golden fixtures of tests/golden/ from the CPU oracle.

Each fixture is 64 columns x 96 levels of one workload: the inputs are the seeded
synthetic columns (regenerated by the tests from the recorded seed, fractions and
offsets, and checked against the recorded input checksum), the outputs are every
array the step writes after `steps` oracle steps at dt = 15 s.

    python tests/golden/make_step_fixtures.py

The fixtures pin the GPU step to the oracle at these inputs and guard the oracle
and the input generator against silent change.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from nicam_kokkos_microphysics_amd import fields as F  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    # 130x130-column profile generator, first 64 columns of region 0
    "plain_1step": dict(seed=20240101, frac_supersat=0.0, frac_towers=0.0, l_offset=0, ij_offset=0, steps=1),
    # the same columns, ten consecutive steps
    "plain_10steps": dict(seed=20240101, frac_supersat=0.0, frac_towers=0.0, l_offset=0, ij_offset=0, steps=10),
    # mixed-precision workload: 5 % supersaturated columns (region 3, columns 4000..4063)
    "supersat_1step": dict(seed=20240101, frac_supersat=0.05, frac_towers=0.0, l_offset=3, ij_offset=4000,
                           steps=1),
    # rollout workload: 10 % deep convective towers (region 17, columns 9000..9063)
    "towers_1step": dict(seed=20240101, frac_supersat=0.0, frac_towers=0.10, l_offset=17, ij_offset=9000,
                         steps=1),
}
NCOL = 64
DT = 15.0


def input_digest(st):
    h = hashlib.sha256()
    for k in F.INPUTS:
        h.update(np.ascontiguousarray(st[k]).tobytes())
    return h.hexdigest()


def make_inputs(case):
    c = CASES[case]
    return F.synth_state(1, NCOL, seed=c["seed"], frac_supersat=c["frac_supersat"], frac_towers=c["frac_towers"],
                         l_offset=c["l_offset"], ij_offset=c["ij_offset"])


def main():
    meta = {}
    for name, c in CASES.items():
        gz, gzh, st = make_inputs(name)
        digest = input_digest(st)
        for _ in range(c["steps"]):
            rep = O.mp_driver(st, gz, gzh, DT)
            assert rep.status == 0, (name, rep.status)
        out = {k: st[k] for k in F.OUTPUTS}
        np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **out)
        meta[name] = dict(c, ncol=NCOL, dt=DT, input_sha256=digest)
    with open(os.path.join(HERE, "step_fixtures.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
