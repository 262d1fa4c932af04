#!/usr/bin/env python3
"""Writes tests/golden/cm1_golden.json by running the reference's own CM1
input_sounding reader and file-sounding column builder: oracle/_ref/ref_sounding,
which `make -C oracle ref` compiles from the reference sources where they lie
(src/soundings/schemes/cm1/cm1_sounding.cpp, src/soundings/base/sounding_algorithms.cpp,
src/soundings/schemes/sharpy/sharpy_interpolation.cpp,
src/init/sounding/sources/file_sounding.cpp and the sounding runtime).

Inputs: the four CM1 soundings of the reference's test fixtures
(tests/fixtures/soundings/cm1/, copied as data to tests/golden/cm1/) and a few
small synthetic texts for the reader's refusals. Outputs (%.17g, exact):
the reader's levels, the model columns for several target grids, every
interpolation method and both extrapolation settings, or the error text.
Run in the development container (the reference tree is not present
elsewhere):

    make -C oracle ref && python tests/golden/make_cm1_golden.py
"""
import json
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_sounding")
SRC = "/root/reference/tests/fixtures/soundings/cm1"
DST = os.path.join(HERE, "cm1")
NAMES = ["HIGHCAPENEW_QUARTERCIRCLE", "HIGHCAPENEW_SEMICIRCLE_02", "MEDCAPENEW_QUARTERCIRCLE",
         "MEDCAPENEW_SEMICIRCLE_06"]

# target columns: (heights, dz)
COLUMNS = [
    ([25.0, 75.0, 125.0, 175.0, 225.0], 50.0),                  # test_file_sounding.cpp:409
    ([250.0 + 500.0 * k for k in range(32)], 500.0),            # test_file_sounding.cpp:565 linspace(250, 15750, 32)
    ([(k + 0.5) * 500.0 for k in range(38)], 500.0),            # 38 x 500 m cells: top interface 19 km
    ([0.0, 3.0, 7.5, 1234.5, 18990.0], 20.0),                   # on a level, between levels, near the top
    ([19500.0], 500.0),                                         # test_file_sounding.cpp:651 exceeds the top
    ([18800.0], 500.0),                                         # :658 upper half-cell uncovered
]

SYNTHETIC = {
    "fortran_exponents_and_comments": "! a comment line\n1.0d3 300.0 10.0  # header\n"
                                      "0 300 10 1 2\n500 301 9 1.5D0 2\n1000 302 8 2 2\n"
                                      "1500 303 7 2 2\n2000 304 6 2 2 ! trailing\n",
    "too_few_levels": "1000 300 10\n0 300 10 0 0\n500 301 9 0 0\n1000 302 8 0 0\n1500 303 7 0 0\n",
    "wrong_column_count": "1000 300 10\n0 300 10 0 0\n500 301 9 0\n1000 302 8 0 0\n1500 303 7 0 0\n"
                          "2000 304 6 0 0\n",
    "not_numeric": "1000 300 10\n0 300 10 0 0\n500 301 x 0 0\n1000 302 8 0 0\n1500 303 7 0 0\n2000 304 6 0 0\n",
    "first_level_not_ground": "1000 300 10\n10 300 10 0 0\n500 301 9 0 0\n1000 302 8 0 0\n1500 303 7 0 0\n"
                              "2000 304 6 0 0\n",
    "heights_not_increasing": "1000 300 10\n0 300 10 0 0\n500 301 9 0 0\n500 302 8 0 0\n1500 303 7 0 0\n"
                              "2000 304 6 0 0\n",
    "header_mismatch": "1000 300.5 10\n0 300 10 0 0\n500 301 9 0 0\n1000 302 8 0 0\n1500 303 7 0 0\n"
                       "2000 304 6 0 0\n",
    "wind_too_large": "1000 300 10\n0 300 10 0 0\n500 301 9 400 0\n1000 302 8 0 0\n1500 303 7 0 0\n"
                      "2000 304 6 0 0\n",
    # one cold level (T < 200 K at 12 km for theta = 300 K) fails the QC bound:
    # the reader refuses rather than drop it
    "qc_drops_a_level": "1000 300 0\n0 300 0 0 0\n1000 300 0 0 0\n2000 300 0 0 0\n3000 300 0 0 0\n"
                        "4000 300 0 0 0\n5000 300 0 0 0\n12000 300 0 0 0\n",
    # most levels cold: fewer than five survive
    "qc_too_few_valid": "1000 300 0\n0 300 0 0 0\n5000 300 0 0 0\n10000 300 0 0 0\n15000 300 0 0 0\n"
                        "20000 300 0 0 0\n25000 300 0 0 0\n",
    # dry constant theta: the Exner function is linear in z (trapezoid rule exact)
    "dry_isentropic": "1000 300 0\n0 300 0 5 -3\n1000 300 0 5 -3\n2000 300 0 5 -3\n3000 300 0 5 -3\n"
                      "4000 300 0 5 -3\n5000 300 0 5 -3\n",
}


def run(args):
    out = subprocess.run([HARNESS] + args, capture_output=True, text=True, check=True).stdout
    rows, error = [], None
    for line in out.splitlines():
        tag, _, rest = line.partition(" ")
        if tag == "error":
            error = rest
        elif tag in ("level", "center"):
            rows.append([float(x) for x in rest.split()])
    return {"rows": rows, "error": error}


def main():
    os.makedirs(DST, exist_ok=True)
    for n in NAMES:
        shutil.copyfile(os.path.join(SRC, n), os.path.join(DST, n))
    gold = {"files": {}, "synthetic": {}}
    for n in NAMES:
        path = os.path.join(DST, n)
        entry = {"read": run(["read", path]), "columns": []}
        for z, dz in COLUMNS:
            for method in (0, 1, 2):
                for extrap in (0, 1):
                    r = run(["column", path, str(method), str(extrap), str(extrap), repr(dz)] + [repr(x) for x in z])
                    entry["columns"].append({"z": z, "dz": dz, "method": method, "extrapolate": extrap, **r})
        gold["files"][n] = entry
    with tempfile.TemporaryDirectory() as d:
        for name, text in SYNTHETIC.items():
            p = os.path.join(d, name)
            with open(p, "w") as f:
                f.write(text)
            e = {"text": text, "read": run(["read", p])}
            if e["read"]["error"] is None:
                e["column"] = run(["column", p, "0", "0", "0", "250", "250", "750", "1250", "1750"])
            gold["synthetic"][name] = e
    with open(os.path.join(HERE, "cm1_golden.json"), "w") as f:
        json.dump(gold, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "cm1_golden.json"))


if __name__ == "__main__":
    main()
