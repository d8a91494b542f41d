"""Writes tests/golden/ref_vectors.json from the reference itself.

* spline_test.F90's known answers (the reference's only unit test), parsed
  from /root/reference/spline_test.F90;
* spline_value of the reference's spline.F90 and hash_double_ of its
  hash_real.c, both compiled from the reference sources into
  oracle/_ref/libref.so by oracle/Makefile, on fixed inputs, as exact
  IEEE-754 bit patterns.

Run from the repository root in the container that holds /root/reference:
    make -f oracle/Makefile && python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import json
import os
import re
import struct
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def known_answers(path="/root/reference/spline_test.F90"):
    out = []
    order = None
    for line in open(path):
        m = re.search(r"spline_init\((\d+)\)", line)
        if m:
            order = int(m.group(1))
        m = re.search(r"assertnear\(([-+0-9.Ee]+),\s*spline_value\(([-+0-9.Ee]+)d0\)\)", line)
        if m:
            out.append({"order": order, "x": float(m.group(2)), "expected": float(m.group(1))})
    return out


def main():
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so missing: run make -f oracle/Makefile")
    rng = np.random.default_rng(1)
    spline = []
    for order in range(2, 9):
        xs = np.concatenate([np.arange(0, order + 0.25, 0.25), rng.uniform(0, order, 40)])
        ref.ref_spline_init(order)
        vals = [ref.ref_spline_value(float(x)) for x in xs]
        ref.ref_spline_cleanup()
        spline.append({"order": order, "x_bits": [bits(x) for x in xs],
                       "value_bits": [bits(v) for v in vals]})
    hashes = []
    for n in (0, 1, 3, 60, 900):
        v = rng.normal(scale=30.0, size=n)
        o = C.c_uint64()
        nn = C.c_int32(n)
        ref.hash_double_(v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nn), C.byref(o))
        hashes.append({"x_bits": [bits(x) for x in v], "hash": int(o.value)})
    doc = {"source": "reference spline.F90, hash_real.c, spline_test.F90",
           "spline_test_known_answers": known_answers(), "spline": spline, "hash": hashes}
    with open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json"), "w") as f:
        json.dump(doc, f, indent=0)


if __name__ == "__main__":
    main()
