"""Binary matrix archives (lv::write / lv::read with lv::MatArchive_BINARY,
modules/utils/src/opencv.cpp:514-531 and :608-625) through the library's host-only entry points.

Pins: three archives written by the reference and held by its own tests (tests/golden/archive/,
copied by tests/golden/copy_archive_fixtures.py) decode to the elements an independent struct-level
parse gives; the LBSP map decodes to the same values as the committed LBSP known-answer fixture
(parsed by tests/golden/make_lbsp_fixture.py, a separate reader); writing a decoded matrix back gives
the reference's bytes exactly; truncated and inconsistent files are refused. No device work."""
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ARCH = os.path.join(ROOT, "tests", "golden", "archive")
FILES = {"test_lbsp.bin": (18, np.uint16, (65, 65, 3)), "test_lss.bin": (5, np.float32, (1, 1, 36)),
         "test_gmm_learn_bg.bin": (6, np.float64, (1, 120))}


def _bgs():
    from litiv_amd import bgs
    return bgs


def _parse(raw):
    typ, esz, cnt, nd = struct.unpack_from("<iQQi", raw, 0)
    sizes = struct.unpack_from("<%di" % nd, raw, 24)
    return typ, esz, cnt, sizes, raw[24 + 4 * nd:]


@pytest.mark.parametrize("name", sorted(FILES))
def test_reference_archives_decode_and_rewrite_byte_identical(name, tmp_path):
    bgs = _bgs()
    typ, dt, shape = FILES[name]
    path = os.path.join(ARCH, name)
    raw = open(path, "rb").read()
    ptyp, esz, cnt, sizes, body = _parse(raw)
    arr, t = bgs.read_archive(path)
    assert t == ptyp == typ and arr.dtype == dt and arr.shape == shape
    assert esz * cnt == len(body) == arr.nbytes
    assert np.array_equal(arr.reshape(-1), np.frombuffer(body, dtype=np.dtype(dt).newbyteorder("<")))
    out = tmp_path / name
    bgs.write_archive(str(out), arr, channels=(typ >> 3) + 1)
    assert out.read_bytes() == raw


def test_lbsp_archive_matches_the_known_answer_fixture():
    bgs = _bgs()
    arr, _ = bgs.read_archive(os.path.join(ARCH, "test_lbsp.bin"))
    exp = np.load(os.path.join(ROOT, "tests", "golden", "lbsp_kat_expected.npy"))
    assert np.array_equal(arr, exp)


def test_header_fields_of_a_written_archive(tmp_path):
    """CV_8UC3 of 2 x 3: type 16, element size 3, 6 elements, 2 dims, then 18 bytes of data."""
    bgs = _bgs()
    a = np.arange(18, dtype=np.uint8).reshape(2, 3, 3)
    p = tmp_path / "a.bin"
    bgs.write_archive(str(p), a, channels=3)
    raw = p.read_bytes()
    assert raw[:32] == struct.pack("<iQQiii", 16, 3, 6, 2, 2, 3)
    assert raw[32:] == a.tobytes() and len(raw) == 50
    b, t = bgs.read_archive(str(p))
    assert t == 16 and np.array_equal(a, b)


def test_truncated_or_inconsistent_archives_are_refused(tmp_path):
    bgs = _bgs()
    raw = open(os.path.join(ARCH, "test_lss.bin"), "rb").read()
    bad = {
        "trunc_data.bin": raw[:-4],
        "trunc_header.bin": raw[:20],
        "count.bin": raw[:12] + struct.pack("<Q", 35) + raw[20:],   # element count != product of sizes
        "esize.bin": raw[:4] + struct.pack("<Q", 8) + raw[12:],     # element size != CV_32FC1's 4
    }
    for n, b in bad.items():
        p = tmp_path / n
        p.write_bytes(b)
        with pytest.raises(bgs.LvbgsError):
            bgs.read_archive(str(p))
    with pytest.raises(bgs.LvbgsError):
        bgs.read_archive(str(tmp_path / "missing.bin"))
