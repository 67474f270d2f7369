"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU; no device):

- the oracle (oracle/oracle.c) built with -fsanitize=address,undefined and driven through its Python binding over
  small cases of every function the parity tests call; its outputs must equal the regular build's;
- the library's host code (JPEG reader, writer, requantisation, trellis, Huffman lengths) in tests/cpp/host_sanitize,
  built by rust-compress_amd/build.py, on the reference's JPEG fixtures and 2,000 deterministic mutations of each
  (truncations, byte flips, random runs, flips inside marker segments).

A sanitizer report ends the process with a non-zero status, which fails the test."""
import glob
import os
import shutil
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE_DIR = os.path.join(ROOT, "oracle")
HOST_BIN = os.path.join(HERE, "cpp", "host_sanitize")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.jpg")) +
                  glob.glob(os.path.join(HERE, "golden", "hier", "*.jpg")))

# Every oracle function the parity tests use, on small inputs; results as arrays in a fixed order.
CASES = textwrap.dedent("""
    import numpy as np
    from oracle import oracle as O

    def run():
        rng = np.random.default_rng(5)
        out = []
        img = rng.integers(0, 256, (29, 37, 3), dtype=np.uint8)
        gray = rng.integers(0, 256, (9, 16), dtype=np.uint8)
        for src, subs in ((img, O.S444), (img, O.S422), (img, O.S420), (gray, O.S420)):
            ci = O.build_coefs(src, 75, subs)
            out += [c.coefs for c in ci.comps]
            out.append(O.reconstruct(ci))
            for trellis in (False, True):
                rq = O.requantize(ci, 50, trellis)
                out += [c.coefs for c in rq.comps]
            plain, scaled = O.build_scaled(src, 60, subs)
            out += list(O.trellis(plain, scaled))
        comp = O.build_coefs(img, 75, O.S420).comps[1]
        bh = len(comp.coefs) // 64 // comp.blocks_w
        plane = O.idct_plane(comp.coefs, comp.blocks_w, bh, np.full(64, 3, np.uint8), 128, 0, 255)
        out.append(plane)
        out.append(O.upsample(plane.astype(np.uint8), 19, 15, 2, 2, 37, 29))
        y, cb, cr, k = (rng.integers(0, 256, 50, dtype=np.uint8) for _ in range(4))
        out.append(O.ycc_to_rgb(y, cb, cr))
        out.append(O.cmyk_to_rgb(y, cb, cr, k, 2))
        a = rng.integers(0, 256, (48, 64, 3), dtype=np.uint8)
        b = O.reconstruct(O.build_coefs(a, 40, O.S420))
        out.append(np.array([O.ssimulacra2(a, b)]))
        q, s, steps = O.find_quality(a, 70.0, 100, O.S420)
        out.append(np.array([q, s, steps], dtype=np.float64))
        out.append(O.build_lengths(np.arange(1, 40, dtype=np.uint32) ** 3, 16))
        return out
    """)


def _regular_results():
    ns = {}
    exec(CASES, ns)
    return ns["run"]()


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], stdout=subprocess.PIPE, text=True, check=True)
    path = p.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    asan = _libasan()
    assert asan, "gcc's AddressSanitizer runtime is not installed"
    lib = str(tmp_path / "liboracle_san.so")
    subprocess.run(["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-mavx2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-shared", "-o", lib, os.path.join(ORACLE_DIR, "oracle.c"), "-lm"], check=True)
    out = str(tmp_path / "san.npz")
    script = CASES + textwrap.dedent(f"""
        res = run()
        np.savez({out!r}, *res)
        print("cases", len(res))
        """)
    env = dict(os.environ, RCG_ORACLE_LIB=lib, LD_PRELOAD=asan, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               PYTHONPATH=ROOT)
    p = subprocess.run([sys.executable, "-c", script], env=env, cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-4000:]
    assert "AddressSanitizer" not in p.stdout and "runtime error" not in p.stdout
    got = np.load(out)
    want = _regular_results()
    assert len(got.files) == len(want)
    for i, w in enumerate(want):
        np.testing.assert_array_equal(got[f"arr_{i}"], np.asarray(w), err_msg=f"case {i}")


def test_host_reader_writer_requantize_trellis_under_sanitizers():
    assert os.path.exists(HOST_BIN), "run __graft_entry__.build() first"
    assert len(FIXTURES) >= 10
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([HOST_BIN, "2000"] + FIXTURES, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=600)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-4000:]
    assert p.stdout.startswith("OK"), p.stdout[-4000:]
    counts = dict(zip(p.stdout.split()[1::2], map(int, p.stdout.split()[2::2])))
    # Each fixture and its mutations ran; the decodable ones went through the writer, requantisation and trellis.
    assert counts["files"] == len(FIXTURES)
    assert counts["decoded"] > counts["files"] and counts["written"] > 0
    assert counts["requantised"] > 0 and counts["trellis"] > 0 and counts["huffman"] > 0
