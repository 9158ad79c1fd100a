"""Writes the golden fixtures of tests/golden from the reference's own translation
units built under oracle/_ref (RngStream.cpp), and copies the data files of the
reference's cylinder physics tests (tests/physics/cyl_160_fine mesh, gzipped; the end
nodes calcE.py measures). Run where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import gzip
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def main():
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libffea_ref.so is not built")
    streams = []
    for seed, stream, subs in [([12345] * 6, 0, 0), ([2024, 77, 5, 190, 2, 99999], 2, 1),
                               ([1, 1, 1, 1, 1, 1], 5, 3)]:
        n = 64
        out = np.zeros(n)
        st = np.zeros(6, np.uint32)
        R.ref_rng_stream_draws(np.array(seed, np.uint32), stream, subs, n, out, st)
        streams.append({"seed": seed, "stream": stream, "substreams": subs,
                        "draws_hex": [float(v).hex() for v in out], "state_after": st.tolist()})
    with open(os.path.join(HERE, "rng_streams.json"), "w") as f:
        json.dump({"source": "reference RngStream.cpp via oracle/_ref", "streams": streams}, f, indent=1)


def cylinder(ref="/root/reference/tests/physics"):
    out = os.path.join(HERE, "cyl_160_fine")
    os.makedirs(out, exist_ok=True)
    for ext in ("node", "top", "surf", "pin"):
        with open(os.path.join(ref, "cyl_160_fine", f"cyl_160_fine.{ext}"), "rb") as fi, \
                open(os.path.join(out, f"cyl_160_fine.{ext}.gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
    src = open(os.path.join(ref, "E_cyl_160_fine", "calcE.py")).read()
    end = [int(x) for x in re.search(r"ENDNODES = \[([^\]]*)\]", src).group(1).split(",")]
    with open(os.path.join(out, "end_nodes.json"), "w") as f:
        json.dump({"source": "tests/physics/E_cyl_160_fine/calcE.py ENDNODES", "end_nodes": end}, f)


if __name__ == "__main__":
    main()
    cylinder()
