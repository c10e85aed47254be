"""Write golden vectors for the per-patch FAC kernels from the reference's own
Fortran (solv_cellpoissonfacops{2d,3d}.f, solv_cartesianrobinbchelper{2d,3d}.f),
compiled from /root/reference by oracle/Makefile into oracle/_ref/.

    make -C oracle ref && python tests/golden/make_golden.py

Output: tests/golden/kernels_ref.npz (arrays, de-duplicated by content) and
tests/golden/kernels_ref.json (case metadata and array keys).  Only data is
stored: seeded inputs and the reference's outputs.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import kernel_cases as kc  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    lib = oracle.ref_lib()
    if lib is None:
        sys.exit("reference kernels not built: make -C oracle ref")
    cases = kc.make_cases()
    store = {}
    meta = []

    def put(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        key = hashlib.sha1(a.tobytes()).hexdigest()[:16]
        store["a" + key] = a
        return "a" + key

    for c in cases:
        out = kc.run_ref(lib, c)
        m = kc.case_to_json(c)
        m["inputs"] = {k: put(v) for k, v in c["arrays"].items()}
        m["expected"] = {k: put(v) for k, v in out.items()}
        meta.append(m)
    np.savez_compressed(os.path.join(HERE, "kernels_ref.npz"), **store)
    with open(os.path.join(HERE, "kernels_ref.json"), "w") as f:
        json.dump({"source": "flang-compiled reference Fortran, oracle/Makefile ref",
                   "cases": meta}, f, indent=0)
    print("wrote %d cases, %d arrays" % (len(meta), len(store)))


if __name__ == "__main__":
    main()
