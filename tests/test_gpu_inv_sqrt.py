"""The stencil weights' 1/sqrt on the GPU (rflu::inv_sqrt_cr, a shortened form of the
compiler's correctly rounded square root and divide) returns the bits of
1.0 / sqrt(x), checked over 2^26 arguments in each of five sets by
tests/native/inv_sqrt_check (built by __graft_entry__.build())."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "native", "inv_sqrt_check")


def test_inv_sqrt_matches_library_bits(gpu):
    if not os.path.exists(EXE):
        pytest.fail("tests/native/inv_sqrt_check is not built (run __graft_entry__.build())")
    print("running", EXE, "26", flush=True)
    p = subprocess.run([EXE, "26"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout
    sets = re.findall(r"^set (\d+): (\d+) arguments, (\d+) mismatches", p.stdout, re.M)
    assert [int(s) for s, _, _ in sets] == [0, 1, 2, 3, 4], p.stdout
    assert all(int(n) == 1 << 26 and int(bad) == 0 for _, n, bad in sets), p.stdout
