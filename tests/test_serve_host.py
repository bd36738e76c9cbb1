"""The serving loop's host code (admission, chunking, round-robin batches,
completion, release, shutdown) under the address, undefined-behaviour and
thread sanitizers, linked against a host model of the HIP runtime
(tests/cpp/fake_hip.cpp) in place of a GPU.  Every returned result is checked
against the pattern the model kernels write for that request."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


@pytest.fixture(scope="module")
def binaries():
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc is needed to build the host-only server objects")
    r = subprocess.run(["make", "-C", CPP, "serve_host", "-j4"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return {k: os.path.join(CPP, "bin", f"serve_host_{k}") for k in ("asan", "tsan")}


def run(path, scale, extra_env):
    env = dict(os.environ)
    env.update(extra_env)
    r = subprocess.run([path, str(scale)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    return r.stdout


def test_server_host_logic_asan_ubsan(binaries):
    out = run(binaries["asan"], 1.0,
              {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
               "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    for line in ("paced: 600 requests ok", "burst: 600 requests ok", "admission:",
                 "oversized and wrapping sizes refused ok", "shutdown with work pending ok"):
        assert line in out, out


def test_server_host_logic_tsan(binaries):
    out = run(binaries["tsan"], 0.5, {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"})
    assert "burst: 300 requests ok" in out and "shutdown with work pending ok" in out, out
