"""Committed oracle output digests (tests/golden/oracle_outputs.json, written by
tests/golden/make_golden.py): the oracle must keep reproducing them, and the HIP library must
produce the same bytes."""
import json
import os

import pytest

import features as F
from golden.make_golden import SCENES, case_outputs, scene_outputs

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_outputs.json")))
CASES = {c.name: c for c in F.CASES}


def test_golden_covers_every_case():
    assert set(GOLD["cases"]) == set(CASES)
    assert set(GOLD["scenes"]) == set(SCENES)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_oracle_scene_matches_golden(cpu_lib, name):
    assert scene_outputs(cpu_lib, name) == GOLD["scenes"][name]


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_case_matches_golden(cpu_lib, name):
    assert case_outputs(cpu_lib, CASES[name]) == GOLD["cases"][name]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCENES))
def test_hip_scene_matches_golden(hip_lib, name):
    assert scene_outputs(hip_lib, name) == GOLD["scenes"][name]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_case_matches_golden(hip_lib, name):
    assert case_outputs(hip_lib, CASES[name]) == GOLD["cases"][name]
