"""The HipComputeBackend adapter (nimbus_amd/adapter/hip_compute_backend.cpp)
compiled against the reference's own headers: it must implement the
reference's ComputeBackend families as they are, link against the C ABI with
no undefined symbol, and behave as a backend on a host without a GPU
(initialize() refuses with a message, families stay unavailable). Skipped
where the reference tree is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
ADAPTER = os.path.join(ROOT, "nimbus_amd", "adapter", "hip_compute_backend.cpp")
DRIVER = os.path.join(ROOT, "tests", "native", "adapter_driver.cpp")
LIB_DIR = os.path.join(ROOT, "nimbus_amd")

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "compute")),
                                reason="the reference tree is not present")

FLAGS = ["-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(REF, "include"),
         "-I" + os.path.join(REF, "src"), "-I" + os.path.join(ROOT, "include")]


def test_adapter_compiles_against_the_reference_headers():
    r = subprocess.run(["g++", "-fsyntax-only"] + FLAGS + [ADAPTER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adapter_links_and_runs_as_a_backend(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "libnimbus_hip.so")):
        pytest.skip("libnimbus_hip.so is not built")
    exe = str(tmp_path / "adapter_driver")
    r = subprocess.run(["g++"] + FLAGS + ["-O1", ADAPTER, DRIVER, "-o", exe, "-L" + LIB_DIR, "-lnimbus_hip",
                                          "-Wl,-rpath," + LIB_DIR, "-Wl,--no-undefined"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert run.stdout.strip().endswith("ok")
