"""The exact parallel sampling table on the host: the block decomposition the
device kernels use (qpp_amd/csrc/qs_exact_core.hpp, run by
tests/cpp/exact_scan_check.cpp) against libstdc++'s own sequential
std::accumulate and std::partial_sum (random.tcc:2656-2712), bit for bit,
from 2^12 to 2^24 elements.

The inputs are chosen to break a wrong rounding rule: exact half-ulp ties
(checked to fail when the tie rule is changed), zero runs and a zero prefix,
one dominant weight, subnormal starts, exact powers of two, and an
approximate block prefix skewed on purpose so that many block maps are
built in the wrong binade (the chain must then fall back to element sums).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "exact_scan_check.cpp")

MODES = ["random", "zeros", "dominant", "ties", "tiny", "powers", "halves"]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("exact") / "exact_scan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-o", exe, SRC],
                   check=True)
    return exe


def run(exe, mode, logn, seed, perturb=0):
    r = subprocess.run([exe, mode, str(logn), str(seed), str(perturb)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("logn", [12, 16, 20])
def test_table_bit_exact(checker, mode, logn):
    run(checker, mode, logn, 100 + logn)
    run(checker, mode, logn, 200 + logn, perturb=1)


@pytest.mark.parametrize("mode", ["random", "zeros", "ties"])
def test_table_bit_exact_2_24(checker, mode):
    out = run(checker, mode, 24, 424, perturb=1)
    # the skewed prefix really sent blocks down the element path
    slow = int(out.split("slow=")[1].split()[0])
    assert slow > 100, out
