"""The host-side schedule of the two-step launch
(tight_binding_amd/csrc/pair_plan.h), compiled with g++ and run by
tests/cpp/pair_schedule.cpp: its invariants on periodic, open, banded and
one-sided (non-symmetric) patterns, and CPU replays of the launch -- the
prologue's band tiles, then workers drawing queue tickets one ahead, each
item reading its inputs once its waits are met and writing at a random later
moment -- equal to two plain steps bit for bit, never stalling."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("pair") / "pair_schedule")
    subprocess.check_call(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall",
                           "-I" + os.path.join(ROOT, "tight_binding_amd", "csrc"), "-o", out,
                           os.path.join(ROOT, "tests", "cpp", "pair_schedule.cpp")])
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pair_schedule_invariants_and_replay(exe, seed):
    r = subprocess.run([exe, str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
