"""Host code under the sanitizers (CPU only): address + undefined-behaviour
sanitizer over the generators, the snapshot and dump I/O and the CPU
oracle; thread sanitizer over concurrent set_jp writes (the reference driver
calls set_jp from OpenMP threads, SRC/nbodysystem.h:753-756).  Built by
tests/sanitize/Makefile into a temporary directory."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "hermite_amd", "libhermite_hip.so")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    if not os.path.exists(LIB):
        pytest.skip("libhermite_hip.so not built")
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-j4", "-C", os.path.join(HERE, "sanitize"), f"OUT={out}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def run(exe, *args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, timeout=600)
    return r.returncode, r.stdout + r.stderr


def test_address_and_ub_sanitizers(built, tmp_path):
    rc, out = run(str(built / "san_io"), str(tmp_path),
                  env_extra={"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1"})
    assert rc == 0, out
    assert "AddressSanitizer" not in out and "runtime error" not in out, out
    assert "0 failed checks" in out


def test_thread_sanitizer_concurrent_set_jp(built):
    rc, out = run(str(built / "san_jstore"), env_extra={"TSAN_OPTIONS": "halt_on_error=1"})
    assert rc == 0, out
    assert "ThreadSanitizer" not in out, out
    assert "0 failed checks" in out
