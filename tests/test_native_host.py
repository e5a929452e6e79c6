"""The host C++ operator layer (fluidity_amd/native: reference-signature DepositMacQuanta3D, GatherMac3D,
ProjectMacPressure3D, SparseMacGrid3D and the substep backend table) through its C++ test program,
which checks it against the oracle. Built by __graft_entry__.build()."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fluidity_amd", "native", "_build", "host_operators_test")
LIB = os.path.join(ROOT, "fluidity_amd", "libfluidity_native.so")


def run(mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build()")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_native_library_loads_and_exports_the_operator_api():
    lib = ctypes.CDLL(LIB)
    # C++ symbols of the reference-shaped API (Itanium mangling of namespace fluidity).
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for name in ("DepositMacQuanta3D", "GatherMac3D", "ProjectMacPressure3D", "HipProjectMacPressure3D",
                 "MakeHipLiquidSubstepBackend3D", "ComputeParticleSortBin3D", "SparseMacGrid3D6Create"):
        assert name in out, name
    assert lib is not None


def test_host_layout_and_status_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("device present: the no-device status test applies to hosts without a GPU")
    run("cpu")


@pytest.mark.gpu
def test_host_operators_match_oracle_on_device():
    run("gpu")
