"""Writes the LBSP known-answer fixture used by the oracle and GPU tests.

Source data (reference repository, read here only to produce the vectors):
  samples/data/multispectral_stereo_ex/img2.png        input image (read by cv::imread as BGR)
  modules/features2d/test/data/test_lbsp.bin           expected CV_16UC3 map (lv::write MatArchive_BINARY)
The reference test (modules/features2d/test/lbsp.cpp:22-46) crops Rect(339,339,65,65) from the
image, runs LBSP(20) (absolute threshold 20) with compute2, and compares the valid zone [2,63)^2.

Outputs (committed): lbsp_kat_input.npy (65x65x3 uint8, BGR) and lbsp_kat_expected.npy
(65x65x3 uint16). Run: python tests/golden/make_lbsp_fixture.py
"""
import os
import struct

import numpy as np
from PIL import Image

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_mat_archive_binary(path):
    # layout: int32 type, uint64 elemSize, uint64 elemCount, int32 dims, int32 sizes[dims], raw data
    with open(path, "rb") as f:
        raw = f.read()
    typ, esz, cnt, dims = struct.unpack_from("<iQQi", raw, 0)
    sizes = struct.unpack_from("<%di" % dims, raw, 24)
    off = 24 + 4 * dims
    depth, ch = typ & 7, (typ >> 3) + 1
    assert depth == 2, "expected CV_16U"
    data = np.frombuffer(raw[off:off + esz * cnt], dtype="<u2").reshape(tuple(sizes) + (ch,))
    return typ, data


def main():
    img = np.asarray(Image.open(os.path.join(REF, "samples/data/multispectral_stereo_ex/img2.png")).convert("RGB"))
    crop = np.ascontiguousarray(img[339:339 + 65, 339:339 + 65, ::-1])  # RGB -> BGR like cv::imread
    typ, exp = read_mat_archive_binary(os.path.join(REF, "modules/features2d/test/data/test_lbsp.bin"))
    assert typ == 18 and exp.shape == (65, 65, 3), (typ, exp.shape)
    np.save(os.path.join(HERE, "lbsp_kat_input.npy"), crop)
    np.save(os.path.join(HERE, "lbsp_kat_expected.npy"), np.ascontiguousarray(exp))
    print("wrote", crop.shape, exp.shape)


if __name__ == "__main__":
    main()
