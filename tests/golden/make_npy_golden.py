#!/usr/bin/env python3
"""Writes tests/golden/npy/*.npy with the reference's own NPY writer
(src/core/output/npy_writer.cpp, built by `make -C oracle ref` into
oracle/_ref/ref_npy): element c in C order is 0.25 c - 1.5. Run in the
development container:

    make -C oracle ref && python tests/golden/make_npy_golden.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_npy")
CASES = [("f4", 3, 4, 5), ("f4", 2, 3), ("f4", 7, 1), ("f8", 2, 3, 4), ("f4", 1, 1, 1), ("f4", 130, 7, 3)]


def main():
    os.makedirs(os.path.join(HERE, "npy"), exist_ok=True)
    for case in CASES:
        name = "_".join(str(x) for x in case) + ".npy"
        subprocess.run([HARNESS, os.path.join(HERE, "npy", name)] + [str(x) for x in case], check=True)
        print("wrote", name)


if __name__ == "__main__":
    main()
