"""Copy the reference's own test data that pins this path into tests/golden/ (run once, in a container that has
the rust-compress tree at /root/reference). Only data files are copied: the SSIMULACRA2 golden pairs and their
expected scores (fixtures/ssim2/, checked by cli/tests/ssimulacra2.rs:10-38) and three hierarchical JPEGs whose
first frame is write_jpeg(build_coefs(...)) at q85 with trellis off (jpeg/src/hierarchical.rs:209-225,
jpeg/examples/make_hier.rs:50-96), plus the baseline JPEG files named below."""
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/fixtures"
DST = os.path.dirname(os.path.abspath(__file__))

for name in sorted(os.listdir(os.path.join(SRC, "ssim2"))):
    if name.endswith(".png") or name == "expected.csv":
        shutil.copyfile(os.path.join(SRC, "ssim2", name), os.path.join(DST, "ssim2", name))
for name in ("hier_dct_huffman.jpg", "hier_gray_dct_then_lossless.jpg", "hier_gray_refine_dct.jpg"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, "hier", name))
# The reference decoder's recorded output for two of them (jpeg/examples/make_hier.rs:66-96, checked sample for sample
# by jpeg/tests/hierarchical.rs:260-291): they pin reconstruct's IDCT, upsampling and colour conversion.
for name in ("hier_dct_huffman.expected.png", "hier_gray_refine_dct.expected.png"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, "hier", name))
# Pictures of jpeg/tests/jpeg.rs:26-59 (PSNR floors) and baseline files of jpeg/tests/jpeg.rs:121-130 (67x45: odd
# chroma sizes), decoded independently by libjpeg through Pillow in the tests.
for name in ("rgb8.png", "gray8.png", "rgb8_420.jpg", "rgb8_444.jpg", "gray8.jpg"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
# Four-component baseline files of cli/tests/cmyk.rs:38-54 (requantisation at q100 changes no coefficient).
# Pillow's RGB conversions of them, the reference's decode tolerance target (cli/tests/cmyk.rs:14-27,58-72).
for name in ("cmyk.jpg", "ycck.jpg", "ycck_444.jpg", "cmyk_expected.png", "ycck_expected.png",
             "ycck_444_expected.png"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
# JPEG sources of the reader (jpeg/tests/jpeg_decode.rs:6,77,126; jpeg/tests/arith.rs:168): baseline, progressive
# (jpegtran-style scan script with successive approximation) and optimised-table files, plus one each of the frame
# types the reader describes but does not decode (arithmetic, 12-bit, lossless, hierarchical progressive).
for name in ("rgb8.jpg", "rgb8_prog.jpg", "rgb8_opt.jpg", "arith_420_seq.jpg", "jpeg12_444.jpg",
             "lossless_gray_p8_pred1.jpg", "hier_dct_progressive.jpg"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
