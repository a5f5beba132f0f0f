"""The oracle (CPU) and the MI355X kernels (GPU) against the committed golden
fixtures of tests/golden (written by tests/golden/make_golden.py)."""
import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as MG  # noqa: E402


def compare(name, runner_cls):
    ref = np.load(os.path.join(os.path.dirname(MG.__file__), f"{name}.npz"))
    out = MG.render_case(name, runner_cls)
    assert set(out) == set(ref.files)
    for k in ref.files:
        H.assert_same(f"{name}:{k}", out[k], ref[k])


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_oracle_matches_golden(name):
    compare(name, H.OracleRunner)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_gpu_matches_golden(gpu, name):
    H.flush(f"golden {name}")
    compare(name, H.GpuRunner)
