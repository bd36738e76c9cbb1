"""The serving loop as it was when it captured its batches into HIP graphs
(commit f589f0a: batches of eight or more launches captured over the worker's
streams in thread-local mode, instantiated, launched and destroyed per batch,
with the AverageOn result copies into pinned slots among the captured nodes),
replayed on the CPU against the HIP model of tests/cpp/fake_hip.cpp, which
models stream capture, under ASan/UBSan and TSan.  That loop ended a
two-server GPU benchmark with a host segmentation fault; the replay drives
several server lifetimes in one process with graph and direct batches mixed
and checks every result.  It passes: the fault is not in the loop's own host
logic as far as the model goes (see the design notes on the serving loop).
Needs hipcc and the git history."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "bin", "replay")


@pytest.fixture(scope="module")
def replay():
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc is needed to build the host-only server objects")
    if shutil.which("git") is None or not os.path.isdir(os.path.join(ROOT, ".git")):
        pytest.skip("the replay builds the loop from the git history")
    ok = subprocess.run(["git", "-C", ROOT, "cat-file", "-e", "f589f0a^{commit}"],
                        capture_output=True)
    if ok.returncode != 0:
        pytest.skip("commit f589f0a is not in this checkout")
    r = subprocess.run(["make", "-C", CPP, "graph_replay"], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return BIN


def run(path, scale, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run([path, str(scale)], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    return r.stdout


def graph_batches(out):
    return sum(int(m) for m in re.findall(r"graph batches (\d+) of", out))


def test_graph_serving_replay_asan_ubsan(replay):
    out = run(os.path.join(replay, "serve_host_asan"), 1.0,
              {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
               "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    assert "paced: 600 requests ok" in out and "burst: 600 requests ok" in out, out
    assert "shutdown with work pending ok" in out, out
    # graphs were captured in more than one server lifetime
    assert len([m for m in re.findall(r"graph batches (\d+) of", out) if int(m) > 0]) >= 2, out
    assert graph_batches(out) > 20, out


def test_graph_serving_replay_tsan(replay):
    out = run(os.path.join(replay, "serve_host_tsan"), 0.5,
              {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"})
    assert "burst: 300 requests ok" in out and "shutdown with work pending ok" in out, out
    assert graph_batches(out) > 10, out
