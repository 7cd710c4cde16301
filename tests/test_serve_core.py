"""CPU checks of the projection server's scheduler (shiokaze_amd/csrc/shkz_serve.hpp): the C++
check program runs many client threads against a stand-in executor, once plainly and once under
ThreadSanitizer (host code only; no GPU involved)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "serve_core_check.cpp")


def build_and_run(tmp_path, flags):
    exe = tmp_path / "serve_core_check"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", *flags, SRC, "-o", str(exe), "-lpthread"])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "OK" in p.stdout, p.stdout + p.stderr
    return p.stdout


def test_scheduler_invariants(tmp_path):
    out = build_and_run(tmp_path, [])
    assert "batches" in out


def test_scheduler_under_thread_sanitizer(tmp_path):
    probe = tmp_path / "probe.cpp"
    probe.write_text("int main(){return 0;}\n")
    if subprocess.run(["g++", "-fsanitize=thread", str(probe), "-o", str(tmp_path / "probe")]).returncode != 0:
        pytest.skip("g++ has no ThreadSanitizer runtime here")
    build_and_run(tmp_path, ["-fsanitize=thread"])
