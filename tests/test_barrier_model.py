"""The group barrier's counting protocol on the CPU (tests/cpp/barrier_model.cpp):
a std::thread model of group_sync with the device's barrier-word layout,
fenced and lean forms, group sizes changing between launches. A planted
mistake (the first workgroup to leave a launch clearing the counters while
others may still wait on them) must make the model fail, so the model is
shown to see that class of bug."""
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


def test_barrier_model_catches_early_clear():
    r = subprocess.run([_build("barrier_model_mutant", "-DCLEAR_AT_FIRST_EXIT"), "60"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout + r.stderr)
    assert r.returncode != 0
