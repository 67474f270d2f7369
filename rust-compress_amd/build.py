"""Build librcgpu.so in-tree with hipcc for gfx950.

One hipcc invocation compiles the HIP kernels and the host C++ into a shared library with no PyTorch dependency.
-ffp-contract=off keeps every f32 operation of the DCT path separately rounded (the SSIMULACRA2 kernels opt back
into contraction per scope); -Xclang -target-feature -Xclang -packed-fp32-ops keeps packed FP32 instructions out of
the generated code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librcgpu.so")
SOURCES = ["api.cpp", "batch.cpp", "server.cpp", "jpeg.hip", "ssim2.hip", "noalloc.hip", "synth.cpp", "writer.cpp",
           "reader.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RCG_OFFLOAD_ARCH", "gfx950")


FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
OBJ_DIR = os.path.join(HERE, "_build")


def command(out=LIB):
    """Single-invocation form of the build (documentation and debugging)."""
    return [HIPCC] + FLAGS + ["-shared", "-o", out] + [os.path.join(CSRC, s) for s in SOURCES] + ["-lpthread"]


ROOT = os.path.dirname(HERE)
API_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "reference_api_test.cpp")
API_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "reference_api_test")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature flag; that note is expected and dropped.
    noise = "is not a recognized feature for this target"
    out = "\n".join(l for l in p.stdout.splitlines() if noise not in l)
    if out.strip():
        print(out, flush=True)
    if p.returncode != 0:
        raise subprocess.CalledProcessError(p.returncode, cmd)


def build(verbose=False):
    """Compile every source to an object in parallel, then link librcgpu.so (and the C++ API test binary)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = [[HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    # C++ tests of the host API (host/rc_gpu.hpp), linked against the library.
    _run(["g++", "-std=c++17", "-O2", "-Wall", API_TEST_SRC, "-o", API_TEST_BIN, "-L" + HERE, "-lrcgpu",
          "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN/../../rust-compress_amd"], verbose)
    build_sanitized(objs, verbose)
    return LIB


# Sanitizer build of the host code: every C++ source compiled again with the host side instrumented by
# AddressSanitizer and UndefinedBehaviorSanitizer (-Xarch_host: device code is never instrumented), linked with the
# kernels' objects of the regular build into two drivers:
#   tests/cpp/host_sanitize      the host reader, writer, requantisation and trellis on the reference's JPEG fixtures
#                                and mutations of them (tests/test_host_sanitizers.py, no device needed);
#   tests/cpp/server_race_asan   the serving queue with concurrent submitting threads (tests/test_gpu_server_race.py).
# The HIP runtime finds no device in a ThreadSanitizer process, so the serving queue's race check under
# ThreadSanitizer runs without one: tests/cpp/server_queue_tsan is the real csrc/server.cpp (and synth.cpp) with the
# same driver, linked against tests/cpp/server_model.cpp, a host model of the device calls and the scoring pass
# (tests/test_server_queue_tsan.py, no device needed).
CLANG = os.environ.get("RCG_CLANGXX", "/opt/rocm/llvm/bin/clang++")
HOST_SOURCES = [s for s in SOURCES if s.endswith(".cpp")]
CHECKS = ["address", "undefined"]
DRIVERS = [(os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"), os.path.join(ROOT, "tests", "cpp", "host_sanitize")),
           (os.path.join(ROOT, "tests", "cpp", "server_race.cpp"),
            os.path.join(ROOT, "tests", "cpp", "server_race_asan"))]


TSAN_SOURCES = [os.path.join(CSRC, "server.cpp"), os.path.join(CSRC, "synth.cpp"),
                os.path.join(ROOT, "tests", "cpp", "server_model.cpp"), os.path.join(ROOT, "tests", "cpp", "server_race.cpp")]
TSAN_BIN = os.path.join(ROOT, "tests", "cpp", "server_queue_tsan")


def build_queue_model(verbose=False):
    """The serving queue and its driver under ThreadSanitizer, over the host model of the device: host code only,
    compiled as C++ against the HIP headers (no HIP runtime linked)."""
    from concurrent.futures import ThreadPoolExecutor
    d = os.path.join(OBJ_DIR, "tsan")
    os.makedirs(d, exist_ok=True)
    flags = [CLANG, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=thread", "-fPIC",
             "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-Wall", "-Wno-unused-result"]
    objs = [os.path.join(d, os.path.splitext(os.path.basename(src))[0] + ".o") for src in TSAN_SOURCES]
    with ThreadPoolExecutor(max_workers=len(objs)) as ex:
        for f in [ex.submit(_run, flags + ["-c", src, "-o", o], verbose) for src, o in zip(TSAN_SOURCES, objs)]:
            f.result()
    _run([CLANG, "-fsanitize=thread", "-o", TSAN_BIN] + objs + ["-lpthread"], verbose)


def _sanitized(kernel_objs, checks, drivers, subdir, verbose):
    from concurrent.futures import ThreadPoolExecutor
    kobjs = [o for o, s in zip(kernel_objs, SOURCES) if not s.endswith(".cpp")]
    flags = [f for f in FLAGS if f != "-O3"] + ["-O1", "-g", "-fno-omit-frame-pointer"]
    for c in checks:
        flags += ["-Xarch_host", "-fsanitize=" + c]
    if "undefined" in checks:
        flags += ["-Xarch_host", "-fno-sanitize-recover=undefined"]
    d = os.path.join(OBJ_DIR, subdir)
    os.makedirs(d, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES] + [drv for drv, _ in drivers]
    objs = [os.path.join(d, os.path.splitext(os.path.basename(src))[0] + ".o") for src in srcs]
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1)) as ex:
        for f in [ex.submit(_run, [HIPCC] + flags + ["-c", src, "-o", o], verbose) for src, o in zip(srcs, objs)]:
            f.result()
    lib_objs = objs[:len(HOST_SOURCES)] + kobjs
    for (drv, out), o in zip(drivers, objs[len(HOST_SOURCES):]):
        _run([CLANG, "-fsanitize=" + ",".join(checks), "-o", out, o] + lib_objs +
             ["-L/opt/rocm/lib", "-lamdhip64", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"], verbose)


def build_sanitized(kernel_objs, verbose=False):
    _sanitized(kernel_objs, CHECKS, DRIVERS, "asan", verbose)
    build_queue_model(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
