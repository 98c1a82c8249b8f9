"""The C++ adapter (aperi-mech_amd/cpp/aperi_hip.hpp) through its own C++ test
program, tests/cpp/test_adapter (built by __graft_entry__.build() / make -C tests/cpp):
material properties, material and element-combination errors on the CPU; the
uniform-shear force identities, the explicit loop and parity with the oracle on the GPU.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_adapter")


def _run(mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build() or make -C tests/cpp")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_adapter_host_cases():
    _run("cpu")


@pytest.mark.gpu
def test_cpp_adapter_device_cases(gpu_available):
    _run("gpu")
