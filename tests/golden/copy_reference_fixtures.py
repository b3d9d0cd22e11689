"""Copies the data files of the reference's demo (test_vfs/) that the fixture tests
use into tests/golden/test_vfs/: the T265 stereo pair of the commented `wang` run
(test_vfs/main.cpp:233-330), its half-resolution mask and vector fields, and the 16-bit
depth and Z images that run wrote (main.cpp:319-324).  Data only: inputs and the
reference's own outputs.

    python tests/golden/copy_reference_fixtures.py [/root/reference]
"""
import os
import shutil
import sys

FILES = ["wang1.png", "wang2.png", "maskHalf.png", "translationVectorWangHalf.flo",
         "calibrationVectorWangHalf.flo", "depth16.png", "z16.png"]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_vfs")
    os.makedirs(dst, exist_ok=True)
    for name in FILES:
        shutil.copyfile(os.path.join(ref, "test_vfs", name), os.path.join(dst, name))
        print("copied", name)


if __name__ == "__main__":
    main()
