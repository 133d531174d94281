"""Runs the C++ host-side test (tests/cpp/test_interfaces.cpp): the reference operator
interfaces re-exposed in C++ (include/hclpp/hydro.hpp) drive the reference main-loop
sequence and the fused driver on the GPU; results must be bit-identical to the oracle."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_interfaces")


def test_cpp_interfaces(cuda):
    assert os.path.exists(BIN), "build it with __graft_entry__.build() (make -C tests/cpp)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
