"""The host code under sanitizers (no GPU): tests/cpp/Makefile builds serve_check.cpp with the serving
loop (hcm_serve.cpp) over a CPU stand-in for the device, the pool allocator, the registry and the
workload generators, once with -fsanitize=address,undefined and once with -fsanitize=thread, and runs
both: every request submitted from four threads completes once, final meshes follow the final blocks,
and neither sanitizer reports anything."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "hypercraft_amd", "_build", "sanitize")


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "-j2", "all"], check=True,
                   timeout=600)


def _run(name):
    p = subprocess.run([os.path.join(OUT, name)], capture_output=True, text=True, timeout=600)
    text = p.stdout + p.stderr
    assert p.returncode == 0, text[-4000:]
    for bad in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "runtime error:", "LeakSanitizer"):
        assert bad not in text, text[-4000:]
    assert "serve_check: all checks passed" in text


def test_host_code_under_address_and_undefined_sanitizers():
    _build()
    _run("serve_check_asan")


def test_host_code_under_thread_sanitizer():
    _build()
    _run("serve_check_tsan")
