"""The oracle's reconstruct, pinned bit for bit against the reference decoder's own recorded output.

hier_dct_huffman.expected.png and hier_gray_refine_dct.expected.png are what rust-compress's decoder produced for
its hierarchical DCT fixtures when they were made (jpeg/examples/make_hier.rs:66-96; checked sample for sample by
jpeg/tests/hierarchical.rs:260-291). Decoding those files runs the f32 dequantisation + AAN IDCT + round_f32 of
idct_planes (unclamped, decode.rs:836-844), the (2,2) fancy upsampling of `upsample` and the fixed-point YCbCr -> RGB
of samples_to_image (decode.rs:882-903) -- the arithmetic of `reconstruct`. tests/jpeg_hier.py does the frame
bookkeeping and calls the oracle for each of those stages, so a mismatch of one sample fails here.

Sensitivity: planted mistakes in a copy of oracle.c (on the CPU) must break the match. Changes of about one ulp do
not change any sample of these small files (about 14,600 IDCT outputs over the frames): a re-association of the
IDCT's f32 operations, or a butterfly constant that rounds to the neighbouring float. Nor does a different tie rule,
since no output lands on a tie; round_f32's ties-to-even is pinned by known answers in test_oracle.py. The IDCT's
exact operation order is therefore pinned by restatement (and by the GPU's bit-exact agreement with it).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from PIL import Image

from oracle import oracle as O
import jpeg_hier

HIER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hier")
ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))
RECORDED = ["hier_dct_huffman", "hier_gray_refine_dct"]


def _recorded(name):
    with open(os.path.join(HIER, name + ".jpg"), "rb") as f:
        data = f.read()
    return data, np.asarray(Image.open(os.path.join(HIER, name + ".expected.png")))


@pytest.mark.parametrize("name", RECORDED)
def test_recorded_hierarchical_decode_is_reproduced_sample_for_sample(name):
    data, want = _recorded(name)
    got = jpeg_hier.decode(data)
    assert got.shape == want.shape
    bad = int(np.count_nonzero(got != want))
    assert bad == 0, f"{name}: {bad} samples differ from the reference decoder's recorded output"


def test_the_colour_fixture_exercises_upsampling_and_colour_conversion():
    # 64x48 picture with 2x2-sampled luma: the chroma planes are 32x24 and go through the (2,2) filter; the result
    # is not grey, so the YCbCr -> RGB constants matter.
    data, want = _recorded("hier_dct_huffman")
    assert want.shape == (48, 64, 3)
    assert np.abs(want[..., 0].astype(int) - want[..., 2].astype(int)).max() > 20


# jpeg/src/hierarchical.rs:447-517: the reference's own unit cases of expand, run on the test-side restatement.
def test_expand_truncates_once_in_the_middle_of_four():
    up = jpeg_hier.expand(np.array([[0, 1], [1, 2]]), True, True)
    assert up[1, 1] == 1
    assert up[0].tolist() == [0, 0, 1, 1]


def test_expand_truncates_toward_zero_and_keeps_the_edges():
    assert jpeg_hier.expand(np.array([[0, 3]]), True, False).tolist() == [[0, 1, 3, 3]]
    assert jpeg_hier.expand(np.array([[10], [21]]), False, True).ravel().tolist() == [10, 15, 21, 21]
    assert jpeg_hier.expand(np.array([[-3, 0]]), True, False).tolist() == [[-3, -1, 0, 0]]


def test_expand_matches_the_per_sample_definition():
    src = np.array([[(x * 37 + y * 91 + x * y * 5) % 251 for x in range(7)] for y in range(5)])
    for hh, vv in [(True, False), (False, True), (True, True), (False, False)]:
        up = jpeg_hier.expand(src, hh, vv)
        assert up.shape == (5 * (1 + vv), 7 * (1 + hh))
        for y in range(up.shape[0]):
            for x in range(up.shape[1]):
                sx, sy = (x // 2 if hh else x), (y // 2 if vv else y)
                pts = [(sx, sy)]
                bx, by = hh and x % 2 == 1, vv and y % 2 == 1
                if bx:
                    pts.append((sx + 1, sy))
                if by:
                    pts.append((sx, sy + 1))
                if bx and by:
                    pts.append((sx + 1, sy + 1))
                s = sum(int(src[min(py, 4), min(px, 6)]) for px, py in pts)
                assert up[y, x] == s // len(pts)


# Planted mistakes, each a plausible slip in restating decode.rs. Every one must change at least one sample.
MUTATIONS = {
    "upsample rounding bias swapped": ("((x % 2 == 0) ? 8 : 7)) >> 4", "((x % 2 == 0) ? 7 : 8)) >> 4"),
    "upsample far column on the wrong side": ("int fi = (x % 2 == 0) ? (i > 0 ? i - 1 : 0)",
                                              "int fi = (x % 2 == 1) ? (i > 0 ? i - 1 : 0)"),
    "transposed dequantisation table": ("mult[v][u] = (float)qt[v * 8 + u] * AAN[u] * AAN[v];",
                                        "mult[v][u] = (float)qt[u * 8 + v] * AAN[u] * AAN[v];"),
    "IDCT odd part sign": ("float o12 = -2.613126f * z10 + z5;", "float o12 = 2.613126f * z10 - z5;"),
    "IDCT even part: (tmp1 + tmp3) for (tmp1 - tmp3)": ("float tmp12 = (tmp1 - tmp3) * SQRT2 - tmp13;",
                                                         "float tmp12 = (tmp1 + tmp3) * SQRT2 - tmp13;"),
    "IDCT final scale left out of one pass": ("b[y][x] *= 0.125f;", "b[y][x] *= 0.25f;"),
    "chroma centred on 127": ("int32_t cb = b - 128, cr = c - 128;", "int32_t cb = b - 127, cr = c - 127;"),
    "green's Cr weight": ("int32_t g = a - ((22554 * cb + 46802 * cr + 32768) >> 16);",
                          "int32_t g = a - ((22554 * cb + 46801 * cr + 32768) >> 16);"),
}


def _mutated_oracle(tmp_path, name, old, new):
    src = open(os.path.join(ORACLE_DIR, "oracle.c")).read()
    assert src.count(old) == 1, f"{name}: the planted text must occur once in oracle.c"
    d = tmp_path / name.replace(" ", "_").replace("'", "")
    d.mkdir()
    (d / "oracle.c").write_text(src.replace(old, new))
    shutil.copy(os.path.join(ORACLE_DIR, "orc_weights.h"), d / "orc_weights.h")
    so = d / "liboracle.so"
    return subprocess.Popen(["gcc", "-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                             "-shared", "-o", str(so), str(d / "oracle.c"), "-lm"]), str(d)


def test_planted_mistakes_in_the_oracle_break_the_recorded_decodes(tmp_path, monkeypatch):
    builds = {name: _mutated_oracle(tmp_path, name, old, new) for name, (old, new) in MUTATIONS.items()}
    files = [_recorded(n) for n in RECORDED]
    undetected = []
    for name, (proc, d) in builds.items():
        assert proc.wait() == 0, name
        monkeypatch.setattr(O, "_HERE", d)
        monkeypatch.setattr(O, "_LIB", None)
        bad = sum(int(np.count_nonzero(jpeg_hier.decode(data) != want)) for data, want in files)
        if bad == 0:
            undetected.append(name)
    monkeypatch.setattr(O, "_LIB", None)
    assert not undetected, f"planted mistakes not seen by the recorded decodes: {undetected}"
