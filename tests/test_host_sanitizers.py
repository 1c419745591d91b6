"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
index builder of every codec (edge lengths, fixed and adaptive hash widths),
the file writer, the refusals of bad input, the synthetic-list generator and
the oracle, each built index decoded and queried back against
std::set_intersection (tests/cpp/host_asan_check.cpp, built by
tests/cpp/Makefile with g++ -fsanitize=address,undefined)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def test_host_code_under_asan_and_ubsan():
    exe = os.path.join(CPP, "host_asan_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", CPP, "host_asan_check"], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failures)" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr
