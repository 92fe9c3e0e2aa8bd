"""Closed-form golden cases (tests/golden/closed_forms.json, written by
tests/golden/make_golden.py from the definition of the operations, not from
any implementation): the oracle's two methods on the CPU, and the HIP
library through its C ABI on the GPU.
"""

import json
import os

import numpy as np
import pytest

import oracle
from voroffset_amd.volume import CompressedVolume

from .helpers import compare_csr

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "closed_forms.json")))["cases"]


def as_volume(case, key):
    n = case["nx"]
    zlo, zhi = case["z_range"]
    return CompressedVolume.from_columns(case["nx"], case["ny"], case[key], origin=(0.0, 0.0, zlo),
                                         extent=(float(n), float(case["ny"]), zhi - zlo), spacing=1.0,
                                         padding=0)


def test_fixture_file_is_what_the_generator_writes(tmp_path):
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(HERE, "golden", "make_golden.py"), str(tmp_path)],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    fresh = open(os.path.join(tmp_path, "closed_forms.json")).read()
    assert fresh == open(os.path.join(HERE, "golden", "closed_forms.json")).read()


@pytest.mark.parametrize("method", ["ours", "brute_force"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_oracle_matches_closed_form(method, case):
    vol = as_volume(case, "input")
    exp = as_volume(case, "expected")
    off, z = oracle.morph(vol, case["op"], case["radius"], method)
    bad, err = compare_csr(off, z, exp.off, exp.z)
    assert len(bad) == 0, f"count mismatch in columns {bad[:8]}"
    assert err <= 1e-12, err


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_hip_fp64_matches_closed_form(hip_fp64, case):
    vol = as_volume(case, "input")
    exp = as_volume(case, "expected")
    res, _, _ = hip_fp64.apply(case["op"], vol, case["radius"])
    bad, err = compare_csr(res.off, res.z, exp.off, exp.z)
    assert len(bad) == 0, f"count mismatch in columns {bad[:8]}"
    assert err <= 1e-9, err


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_hip_fp32_matches_closed_form(hip_fp32, case):
    vol = as_volume(case, "input")
    exp = as_volume(case, "expected")
    res, _, _ = hip_fp32.apply(case["op"], vol, case["radius"])
    bad, err = compare_csr(res.off, res.z, exp.off, exp.z)
    assert len(bad) == 0, f"count mismatch in columns {bad[:8]}"
    assert err <= 1e-4, err
