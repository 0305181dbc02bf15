"""Copies the input files the reference ships in its in/ directory (the
sample params.in deck and the small solids images 3x3.bmp, 9x9.bmp,
3x3x3.raw, 9x9x9.raw) into tests/golden/ref_in as data fixtures for the
deck-reader tests.  Run once where the reference checkout is present:

    python tests/golden/copy_reference_inputs.py /path/to/reference
"""
import pathlib
import shutil
import sys

FILES = ["params.in", "3x3.bmp", "9x9.bmp", "3x3x3.raw", "9x9x9.raw"]


def main(ref):
    src = pathlib.Path(ref) / "in"
    dst = pathlib.Path(__file__).resolve().parent / "ref_in"
    dst.mkdir(exist_ok=True)
    for name in FILES:
        shutil.copyfile(src / name, dst / name)
        print("copied", name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
