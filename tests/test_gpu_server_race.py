"""The serving queue under concurrent callers on the GPU: tests/cpp/server_race.cpp (built by __graft_entry__.build()
with the host side of the library instrumented by AddressSanitizer and UndefinedBehaviorSanitizer) runs four
submitting threads, the polling main thread and a statistics reader against one server, with requests in both lanes.
Every request completes exactly once with its alone result, and no sanitizer report ends the process. The same
driver runs under ThreadSanitizer over a host model of the device in tests/test_server_queue_tsan.py."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")


def _run(binary, env_extra):
    path = os.path.join(HERE, binary)
    assert os.path.exists(path), "run __graft_entry__.build() first"
    env = dict(os.environ, **env_extra)
    print(f"{binary}: 48 requests from 4 threads", flush=True)
    p = subprocess.run([path], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(p.stdout[-6000:])
    assert "Sanitizer" not in p.stdout and "runtime error" not in p.stdout, p.stdout[-6000:]
    assert p.returncode == 0, p.stdout[-6000:]
    assert p.stdout.strip().splitlines()[-1].startswith("OK")


def test_server_queue_with_concurrent_submitters():
    _run("server_race_asan", {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
                              "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})

