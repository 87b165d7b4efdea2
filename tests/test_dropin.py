"""Builds tests/cpp/test_dropin.cpp against include/planetsim/ and libpsim.so
and runs its host part here and its device part on a GPU."""
import ctypes
import os
import subprocess
import sys

import pytest

import planet_game_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def _build(target, extra):
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")
    libdir = os.path.dirname(P.LIB_PATH)
    if not os.path.exists(target) or os.path.getmtime(target) < max(
            os.path.getmtime(src), os.path.getmtime(P.LIB_PATH),
            *(os.path.getmtime(os.path.join(ROOT, "include", "planetsim", f))
              for f in os.listdir(os.path.join(ROOT, "include", "planetsim")))):
        subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-ffp-contract=off",
                        "-I", os.path.join(ROOT, "include"), *extra, src, P.LIB_PATH, "-o", target,
                        f"-Wl,-rpath,{libdir}"], check=True)
    return target


@pytest.fixture(scope="module")
def dropin_exe():
    return _build(os.path.join(BUILD, "test_dropin"), [])


@pytest.fixture(scope="module")
def dropin_library():
    return _build(os.path.join(BUILD, "libtest_dropin.so"), ["-DDROPIN_LIBRARY", "-shared", "-fPIC"])


def test_dropin_host_part(dropin_exe):
    r = subprocess.run([dropin_exe, "cpu"], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_dropin_device_part(dropin_library):
    """In this process: the solver's multi-workgroup launches are
    cooperative, and a second process's cooperative launch on a GPU that
    this process already runs them on is refused (hipErrorIllegalState)."""
    print("[gpu case] C++ drop-in device part", flush=True)
    lib = ctypes.CDLL(dropin_library)
    lib.dropin_run.restype = ctypes.c_int
    lib.dropin_run.argtypes = [ctypes.c_char_p]
    assert lib.dropin_run(b"gpu") == 0
