"""A C program compiled against include/srx.h (no ctypes) renders the plain small scene through the
C ABI -- map/unmap, state getters and a query polled with do_not_wait included -- and its bytes
must equal the committed golden digest of the oracle's render of the same scene."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_outputs.json")))


def build_and_run(lib_path, tmp_path):
    exe = str(tmp_path / "plain_scene")
    lib_dir, lib_name = os.path.split(lib_path)
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "plain_scene.c"), "-o", exe, "-L", lib_dir, f"-l:{lib_name}",
           f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    out = str(tmp_path / "out.bin")
    res = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stderr
    return hashlib.sha256(open(out, "rb").read()).hexdigest()


def test_c_program_on_oracle(cpu_lib, tmp_path):
    assert build_and_run(cpu_lib.path, tmp_path) == GOLD["scenes"]["plain_small_512"]["color_ds"]


@pytest.mark.gpu
def test_c_program_on_hip(hip_lib, tmp_path):
    assert build_and_run(hip_lib.path, tmp_path) == GOLD["scenes"]["plain_small_512"]["color_ds"]


def build_and_run_sasl(lib_path, tmp_path):
    exe = str(tmp_path / "sasl_quad")
    lib_dir, lib_name = os.path.split(lib_path)
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "sasl_quad.c"), "-o", exe, "-L", lib_dir, f"-l:{lib_name}",
           f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, (res.returncode, res.stderr)


def test_c_sasl_program_on_oracle(cpu_lib, tmp_path):
    """A SASL pixel shader compiled and run on one quad through the C ABI (oracle)."""
    build_and_run_sasl(cpu_lib.path, tmp_path)


@pytest.mark.gpu
def test_c_sasl_program_on_hip(hip_lib, tmp_path):
    build_and_run_sasl(hip_lib.path, tmp_path)
