"""The step's square roots (swe::sqrt_rn, swe::sqrt_normal of csrc/swe_kernels.h) against the device
sqrt(), bit for bit, on 2^26 hashed arguments of each class: normal arguments for sqrt_normal, and
every bit pattern (zeros, negatives, denormals, infinities, NaNs) and tiny arguments for sqrt_rn.
tests/hip/sqrt_check is built by __graft_entry__.build() (make -C tests/hip)."""
import os
import subprocess

import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

CHECK = os.path.join(REPO, "tests", "hip", "_build", "sqrt_check")


@pytest.mark.skipif(not os.path.exists(CHECK), reason="tests/hip/_build/sqrt_check is not built (make -C tests/hip)")
def test_square_roots_match_sqrt_bit_for_bit():
    r = subprocess.run([CHECK, "26"], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("class")]
    assert len(lines) == 5 and all(" 0 of " in l for l in lines)
