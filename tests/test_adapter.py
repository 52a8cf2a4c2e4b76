"""The Parthenon-side adapter (pangu_amd/adapter/pangu_hip_tasks.hpp: the
task functions of ideal_grmhd.cc:62-110 over the C ABI) compiles against
the interface it uses, links against libpangu_hip.so and maps an input
file's blocks to a pangu_config (CPU; runs no device code)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adapter_compiles_links_and_maps_the_input(tmp_path):
    gxx = shutil.which("g++")
    lib = os.path.join(ROOT, "pangu_amd", "libpangu_hip.so")
    if gxx is None or not os.path.exists(lib):
        pytest.skip("g++ or libpangu_hip.so missing")
    exe = str(tmp_path / "adapter_compile")
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Werror",
           "-I", os.path.join(ROOT, "tests", "adapter", "mock"),
           "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "pangu_amd", "adapter"),
           os.path.join(ROOT, "tests", "adapter", "adapter_compile.cpp"),
           "-o", exe, lib, "-Wl,-rpath," + os.path.dirname(lib)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "config ok" in out.stdout and "unknown-choice rejected" in out.stdout
