"""The group barrier's counting protocol on the CPU (tests/cpp/barrier_model.cpp):
a std::thread model of group_sync with the device's barrier-word layout,
fenced and lean forms, group sizes changing between launches. A planted
mistake (the last arrival clearing only its own leaf of the idle set) must
make the model fail, so the model is shown to see that class of bug."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "barrier_model.cpp")
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def _build(name, *defines):
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, name)
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-pthread", *defines, SRC, "-o", exe], check=True)
    return exe


def test_barrier_protocol_model():
    r = subprocess.run([_build("barrier_model"), "60"], capture_output=True, text=True, timeout=300)
    print(r.stdout + r.stderr)
    assert r.returncode == 0
    assert "fenced form: ok" in r.stdout and "lean form: ok" in r.stdout


def test_barrier_model_catches_partial_clear():
    r = subprocess.run([_build("barrier_model_mutant", "-DCLEAR_OWN_LEAF_ONLY"), "60"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout + r.stderr)
    assert r.returncode != 0
