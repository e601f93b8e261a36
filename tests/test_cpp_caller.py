"""A C++ program compiled against include/kadath_amd.hpp (the System_of_eqs /
Space_spheric / Scalar front end over the C-ABI) runs tutorials/tuto5.cpp's steps.

CPU: it compiles with the host compiler, links libkadath_amd.so, declares the system
and stops cleanly at the first device call (exit status 3).  GPU: it reaches the
reference's known answer 1 + a/r within 1e-8 at every collocation point, with the
single LU and with the domain-decomposed solver (NEWTON_F64_DD), and a declaration with
an operator that has no device form is rejected (exit status 0).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kadath_frontera_amd")


def _compile(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "tuto5_caller")
    cmd = [cxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "tuto5_caller.cpp"), "-L", PKG, "-lkadath_amd",
           "-Wl,-rpath," + PKG, "-o", exe]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def _gpu_present():
    import torch
    return torch.cuda.device_count() > 0


def test_cpp_caller_compiles_and_stops_cleanly_without_device(tmp_path):
    if _gpu_present():
        pytest.skip("a GPU is present: the GPU test runs the program")
    exe = _compile(tmp_path)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert p.returncode == 3, p.stdout + p.stderr
    assert "inner radius of the first shell is 1" in p.stdout


@pytest.mark.gpu
def test_cpp_caller_reaches_tuto5_known_answer(tmp_path):
    exe = _compile(tmp_path)
    print("tuto5 caller: running", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "rejected:" in p.stdout
    assert "domain-decomposed: max |numerical - analytic|" in p.stdout
