"""The continuous-batching server (n-body_amd/csrc/serve.cpp) built with
ThreadSanitizer (and, separately, AddressSanitizer and
UndefinedBehaviorSanitizer) and driven by tests/native/serve_check.cpp: client threads
submit, poll and wait on jobs of both lanes while a thread reads the
statistics, some jobs turn non-finite and must fail alone, every other
result is checked bit for bit against the same steps applied alone, and the
server is destroyed with a client still blocked in its wait.  The forest
calls are served by a host-memory stand-in inside the driver, so the
server's own threading runs unchanged without a GPU.

Built with ROCm's clang: the ThreadSanitizer runtime of the image's g++ 11
does not intercept pthread_cond_clockwait, which std::condition_variable's
timed wait calls, and then reports a false double lock."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _clang():
    for c in ("/opt/rocm/lib/llvm/bin/clang++", "/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_clang() is None, reason="needs clang++")
def test_server_under_thread_sanitizer(tmp_path):
    exe = tmp_path / "serve_check"
    subprocess.run([_clang(), "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "n-body_amd", "csrc"),
                    os.path.join(ROOT, "n-body_amd", "csrc", "serve.cpp"),
                    os.path.join(ROOT, "tests", "native", "serve_check.cpp"), "-o", str(exe), "-lpthread"],
                   check=True, capture_output=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "all checks passed" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_server_under_address_and_ub_sanitizers(tmp_path):
    exe = tmp_path / "serve_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "n-body_amd", "csrc"),
                    os.path.join(ROOT, "n-body_amd", "csrc", "serve.cpp"),
                    os.path.join(ROOT, "tests", "native", "serve_check.cpp"), "-o", str(exe), "-lpthread"],
                   check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "all checks passed" in r.stdout
