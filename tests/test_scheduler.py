"""Host-side checks of the serving loop: the scheduler policy (admission, FIFO, budget,
chunking, cancel, drain, concurrent submitters) through a C++ harness with a fake executor
(tests/cpp/sched_test.cpp), and the spike-stream digest helper."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from spice_amd.serve import mix64, spike_digest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("sanitizers", ["address,undefined", "thread"])
def test_scheduler_policy(tmp_path, sanitizers):
    """The policy harness built with AddressSanitizer + UndefinedBehaviorSanitizer, and with
    ThreadSanitizer (the worker thread, submitters, waiters and cancel share the queue under a
    mutex and condition variable: a data race fails the run). The ThreadSanitizer build uses
    LLVM's clang: GCC 11's runtime does not intercept pthread_cond_clockwait, which libstdc++
    uses for condition_variable::wait_for on steady_clock, and then reports a false "double
    lock" in the next wait."""
    exe = str(tmp_path / "sched_test")
    cxx = "g++"
    if sanitizers == "thread":
        cxx = next((c for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")) if c and os.path.exists(c)), None)
        if cxx is None:
            pytest.skip("ThreadSanitizer build needs clang++")
    subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=" + sanitizers,
                    "-I" + os.path.join(ROOT, "spice_amd", "csrc"), os.path.join(ROOT, "tests", "cpp", "sched_test.cpp"),
                    "-o", exe], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr


def test_mix64_is_the_splitmix64_finalizer():
    # splitmix64 seeded with 0: the first output is mix64(0x9E3779B97F4A7C15) = 0xE220A8397B1DCDAF
    assert int(mix64([0x9E3779B97F4A7C15])[0]) == 0xE220A8397B1DCDAF
    assert int(mix64([0])[0]) == 0


def test_spike_digest_is_order_independent_and_step_sensitive():
    rng = np.random.default_rng(0)
    steps = [(t, rng.choice(1000, size=int(rng.integers(0, 40)), replace=False)) for t in range(50)]
    d = spike_digest(steps)
    assert d == spike_digest([(t, ids[::-1]) for t, ids in steps])
    moved = [(t + (1 if t == 10 else 0), ids) for t, ids in steps]
    assert any(len(ids) for t, ids in steps if t == 10) and spike_digest(moved) != d
    assert spike_digest([]) == 0
