"""The FP64 division of the hydro kernels (physics.hpp ``fdiv``: the hardware division sequence
without its range-scaling instructions) must equal the IEEE quotient bit for bit on its stated
range.  ``tests/cpp/div_check`` divides 2^24 random and adversarial operand pairs on the GPU and
compares every in-range quotient with the host's IEEE division; out-of-range pairs are only
reported, those with a subnormal divisor apart (fdiv gives NaN there, as its header states).
The slope limiters, whose divisors reach the subnormal range in a decayed far field, divide with
IEEE division: on 2^22 (diffR, diffL) pairs, many with a subnormal diffL, an overflowing
quotient or an overflowing square of it, all four limiters must equal the host's evaluation bit
for bit, so no NaN appears where the reference's slope is finite."""

import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "div_check")


def test_kernel_division_is_ieee_on_its_range(cuda):
    assert os.path.exists(BIN), "build it with __graft_entry__.build() (make -C tests/cpp)"
    print("[division] running tests/cpp/div_check", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"in-range: (\d+) pairs, (\d+) mismatches", r.stdout)
    assert m and int(m.group(1)) > 15_000_000 and int(m.group(2)) == 0, r.stdout
    m = re.search(r"subnormal divisors: (\d+) pairs", r.stdout)
    assert m and int(m.group(1)) > 100_000, r.stdout
    m = re.search(r"limiters: (\d+) pairs, (\d+) mismatches \((\d+) with a subnormal diffL", r.stdout)
    assert m and int(m.group(1)) == 1 << 22 and int(m.group(2)) == 0 and int(m.group(3)) > 500_000, r.stdout
