"""The serving queue under concurrent callers on the GPU: tests/cpp/server_race_asan (built by
__graft_entry__.build() with the host side of the library instrumented by AddressSanitizer and
UndefinedBehaviorSanitizer) runs four submitting threads, the polling main thread and a statistics reader against one
server. Every request completes exactly once with its alone result, and no sanitizer report ends the process."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "server_race_asan")


def test_server_queue_with_concurrent_submitters():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    print("server_race_asan: 48 requests from 4 threads", flush=True)
    p = subprocess.run([BIN], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(p.stdout[-6000:])
    assert "Sanitizer" not in p.stdout and "runtime error" not in p.stdout, p.stdout[-6000:]
    assert p.returncode == 0, p.stdout[-6000:]
    assert p.stdout.strip().splitlines()[-1].startswith("OK")
