"""Builds tests/cpp/test_dropin.cpp against include/planetsim/ and libpsim.so
and runs its host part here and its device part on a GPU."""
import os
import subprocess
import sys

import pytest

import planet_game_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


@pytest.fixture(scope="module")
def dropin_exe():
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "test_dropin")
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    libdir = os.path.dirname(P.LIB_PATH)
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(
            os.path.getmtime(src), os.path.getmtime(P.LIB_PATH),
            *(os.path.getmtime(os.path.join(ROOT, "include", "planetsim", f))
              for f in os.listdir(os.path.join(ROOT, "include", "planetsim")))):
        subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "include"), src, P.LIB_PATH, "-o", exe,
                        f"-Wl,-rpath,{libdir}"], check=True)
    return exe


def test_dropin_host_part(dropin_exe):
    r = subprocess.run([dropin_exe, "cpu"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_dropin_device_part(dropin_exe):
    print("[gpu case] C++ drop-in device part", flush=True)
    r = subprocess.run([dropin_exe, "gpu"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr
