"""Whole-step golden fixtures (tests/golden/step_*.npz, written by
tests/golden/make_step_fixtures.py from the CPU oracle): 64 columns x 96 levels of
each workload's generator, every output array after 1 or 10 steps at dt = 15 s.

* CPU: the generator still produces the recorded inputs (SHA-256 of every input
  array) and the oracle still reproduces the recorded outputs (to 1e-13 of each
  field's max-abs), so neither drifts silently.
* GPU: the HIP step reproduces the recorded outputs within 1e-10 of each field's
  max-abs (the FP64 parity bar).
"""
import json
import os
import sys

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd.check import compare_states

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_step_fixtures as MF  # noqa: E402

with open(os.path.join(HERE, "golden", "step_fixtures.json")) as fh:
    META = json.load(fh)
CASES = sorted(META)
KEYS = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
    [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]


def _fixture(case):
    with np.load(os.path.join(HERE, "golden", f"step_{case}.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", CASES)
def test_generator_reproduces_fixture_inputs(case):
    _, _, st = MF.make_inputs(case)
    assert MF.input_digest(st) == META[case]["input_sha256"]


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_fixture_outputs(case):
    from oracle import oracle as O
    gz, gzh, st = MF.make_inputs(case)
    for _ in range(META[case]["steps"]):
        assert O.mp_driver(st, gz, gzh, META[case]["dt"]).status == 0
    bad = compare_states(st, _fixture(case), KEYS, 1e-13)
    assert not bad, bad


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_reproduces_fixture_outputs(case):
    import torch
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = MF.make_inputs(case)
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=META[case]["ncol"])
    print(f"golden {case}", flush=True)
    for _ in range(META[case]["steps"]):
        drv.step(dev, META[case]["dt"])
    torch.cuda.synchronize()
    out = F.to_host(dev)
    drv.close()
    bad = compare_states(out, _fixture(case), KEYS, 1e-10)
    assert not bad, bad
