"""Copies three binary matrix archives that the reference's own tests hold into tests/golden/archive/.

They are data files written by lv::write(..., lv::MatArchive_BINARY)
(modules/utils/src/opencv.cpp:514-531) and read by the reference's tests with lv::read:
  modules/features2d/test/data/test_lbsp.bin        CV_16UC3, 65 x 65       (lbsp.cpp:37)
  modules/features2d/test/data/test_lss.bin         CV_32FC1, 1 x 1 x 36    (lss.cpp:44)
  modules/imgproc/test/data/test_gmm_learn_bg.bin   CV_64FC1, 1 x 120       (imgproc/test/global.cpp:255)
The archive tests read them with the library's reader, check the decoded elements against an
independent parse, and check that writing the decoded matrix back gives the same bytes.
Run: python tests/golden/copy_archive_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "archive")
FILES = [
    "modules/features2d/test/data/test_lbsp.bin",
    "modules/features2d/test/data/test_lss.bin",
    "modules/imgproc/test/data/test_gmm_learn_bg.bin",
]


def main():
    os.makedirs(HERE, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(REF, f), os.path.join(HERE, os.path.basename(f)))
        print("copied", f)


if __name__ == "__main__":
    main()
