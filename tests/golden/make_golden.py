"""Regenerates tests/golden/ref_fixtures.json from the reference's own compilable sources.

Builds oracle/_ref/ref_fixtures (oracle/Makefile target `ref`, which compiles
LegendrePolynomial_t.cpp, Element2D_t.cpp, Face2D_t.cpp, Hilbert.cpp and Hilbert_splitting.cpp
where they lie under /root/reference together with oracle/ref_fixtures.cpp) and stores its
output. Needs the reference checkout; the committed JSON is what the tests read.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_fixtures")], check=True,
                         capture_output=True, text=True).stdout
    data = json.loads(out)
    with open(os.path.join(HERE, "ref_fixtures.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
    print("wrote", len(out), "bytes of fixtures;", ", ".join(sorted(data)))


if __name__ == "__main__":
    sys.exit(main())
