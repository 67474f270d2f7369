"""The serving queue (rust-compress_amd/csrc/server.cpp, unchanged) under ThreadSanitizer, without a device: the
driver of tests/cpp/server_race.cpp (four submitting threads, the polling main thread and a statistics reader against
one server, requests in both lanes) linked against tests/cpp/server_model.cpp, a host model of the device calls and
the scoring pass whose scores depend only on the uploaded source and the quality. Every request completes exactly
once, requests of the same image and target get the same result, and ThreadSanitizer reports no data race or lock
misuse in the submission, admission, tick, block-reuse, completion and shutdown paths. (The HIP runtime finds no
device inside a ThreadSanitizer process, so this is the form that can run; the same driver runs on the GPU under
AddressSanitizer in tests/test_gpu_server_race.py.)"""
import os
import subprocess

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "server_queue_tsan")


def test_server_queue_under_thread_sanitizer():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    p = subprocess.run([BIN], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(p.stdout[-6000:])
    assert "ThreadSanitizer" not in p.stdout, p.stdout[-6000:]
    assert p.returncode == 0, p.stdout[-6000:]
    last = p.stdout.strip().splitlines()[-1]
    assert last.startswith("OK 48 requests, 21 distinct searches"), last
