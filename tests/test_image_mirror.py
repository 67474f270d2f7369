"""The C++ mirror's Image (rust-compress_amd/host/rc_gpu.hpp: Image::create, to_depth8, flatten; core/src/image.rs:
104-190) on the CPU, against the command-line driver's own flattening (cli._blend and decode_png, written
separately in Python) and the closed forms the reference's formulas fix: a fully opaque sample keeps its value, a
fully transparent one takes the background, and 16-bit 0 and 65535 map to 0 and 255. tests/cpp/image_mirror.cpp is
compiled here with g++; it touches no device."""
import importlib
import io
import os
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "image_mirror.cpp")
cli = importlib.import_module("rust-compress_amd.cli")

GRAY, GRAY_ALPHA, RGB, RGBA = 0, 1, 2, 3


@pytest.fixture(scope="module")
def mirror(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mirror") / "image_mirror")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-o", exe, SRC], check=True)
    return exe


def _flatten(exe, w, h, color, depth, data, bg):
    hdr = struct.pack("<4I3B", w, h, color, depth, *bg)
    p = subprocess.run([exe], input=hdr + bytes(data), stdout=subprocess.PIPE, check=True, timeout=60)
    status = struct.unpack_from("<I", p.stdout)[0]
    if status:
        return status, None
    ow, oh, oc, od = struct.unpack_from("<4I", p.stdout, 4)
    ch = 1 if oc == GRAY else 3
    assert (ow, oh, od) == (w, h, 8) and oc in (GRAY, RGB)
    return 0, np.frombuffer(p.stdout[20:], np.uint8).reshape(h, w, ch)


def test_rgba_and_gray_alpha_flatten_as_the_driver_does(mirror):
    rng = np.random.default_rng(11)
    bg = (10, 200, 255)
    rgba = rng.integers(0, 256, size=(9, 11, 4), dtype=np.uint8)
    st, got = _flatten(mirror, 11, 9, RGBA, 8, rgba.tobytes(), bg)
    assert st == 0
    want = np.stack([cli._blend(rgba[:, :, c], rgba[:, :, 3], bg[c]) for c in range(3)], axis=2)
    np.testing.assert_array_equal(got, want)
    ga = rng.integers(0, 256, size=(5, 7, 2), dtype=np.uint8)
    st, got = _flatten(mirror, 7, 5, GRAY_ALPHA, 8, ga.tobytes(), bg)
    assert st == 0
    np.testing.assert_array_equal(got[:, :, 0], cli._blend(ga[:, :, 0], ga[:, :, 1], bg[0]))


def test_every_sample_and_alpha_pair(mirror):
    # All 65,536 (sample, alpha) pairs, three backgrounds at once (one per channel): a rounding or weighting slip in
    # either implementation shows on some pair.
    c, a = np.meshgrid(np.arange(256, dtype=np.uint8), np.arange(256, dtype=np.uint8))
    rgba = np.stack([c, 255 - c, c, a], axis=2)
    bg = (0, 97, 255)
    st, got = _flatten(mirror, 256, 256, RGBA, 8, rgba.tobytes(), bg)
    assert st == 0
    want = np.stack([cli._blend(rgba[:, :, k], a, bg[k]) for k in range(3)], axis=2)
    np.testing.assert_array_equal(got, want)


def test_every_sample_and_alpha_pair_at_the_extremes(mirror):
    # All 256 samples with alpha 255 keep their value; with alpha 0 they all take the background.
    c = np.arange(256, dtype=np.uint8)
    for a, bg, want in ((255, (3, 4, 5), np.stack([c, c, c], 1)), (0, (3, 4, 5), np.tile([3, 4, 5], (256, 1)))):
        rgba = np.stack([c, c, c, np.full(256, a, np.uint8)], 1)
        st, got = _flatten(mirror, 256, 1, RGBA, 8, rgba.tobytes(), bg)
        assert st == 0
        np.testing.assert_array_equal(got.reshape(256, 3), want)


def test_opaque_images_are_unchanged(mirror):
    rng = np.random.default_rng(12)
    for color, ch in ((GRAY, 1), (RGB, 3)):
        px = rng.integers(0, 256, size=(6, 13, ch), dtype=np.uint8)
        st, got = _flatten(mirror, 13, 6, color, 8, px.tobytes(), (0, 0, 0))
        assert st == 0
        np.testing.assert_array_equal(got, px)


def test_sixteen_bit_samples_round_as_the_driver_does(mirror):
    from PIL import Image
    v = np.array([[0, 1, 128, 129, 32767, 32768, 32896, 65534, 65535]], np.uint16)
    st, got = _flatten(mirror, v.shape[1], 1, GRAY, 16, v.astype(">u2").tobytes(), (255, 255, 255))
    assert st == 0
    buf = io.BytesIO()
    Image.fromarray(v).save(buf, "PNG")
    np.testing.assert_array_equal(got[:, :, 0], cli.decode_png(buf.getvalue(), (255, 255, 255)))
    assert got[0, 0, 0] == 0 and got[0, -1, 0] == 255
    # 16-bit RGBA: each sample brought to 8 bits first, then blended (flatten calls to_depth8 first).
    rng = np.random.default_rng(13)
    rgba16 = rng.integers(0, 65536, size=(4, 5, 4), dtype=np.uint16)
    st, got = _flatten(mirror, 5, 4, RGBA, 16, rgba16.astype(">u2").tobytes(), (1, 2, 3))
    assert st == 0
    d8 = ((rgba16.astype(np.uint32) * 255 + 32767) // 65535).astype(np.uint8)
    want = np.stack([cli._blend(d8[:, :, c], d8[:, :, 3], (1, 2, 3)[c]) for c in range(3)], axis=2)
    np.testing.assert_array_equal(got, want)


def test_create_rejects_bad_depth_and_length(mirror):
    # core/src/image.rs:116-124: depth other than 8 or 16 is Unsupported (status 2), a wrong length Invalid (1).
    assert _flatten(mirror, 2, 2, RGB, 12, bytes(12), (0, 0, 0))[0] == 2
    assert _flatten(mirror, 2, 2, RGB, 8, bytes(11), (0, 0, 0))[0] == 1
    assert _flatten(mirror, 2, 2, RGBA, 16, bytes(31), (0, 0, 0))[0] == 1
