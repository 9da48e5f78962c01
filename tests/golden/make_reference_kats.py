"""Extracts the known-answer vectors of the reference's inverse_quad_mapping_test.cu (three quads,
each with 25 reference/physical coordinate pairs) into tests/golden/reference_kats.json.

The vectors are data held by the reference's own test (tests/inverse_quad_mapping_test.cu:9, :81,
:154); this script reads them as text from the reference checkout. The committed JSON is what the
tests read.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/inverse_quad_mapping_test.cu"


def vectors(block):
    return [[float(a), float(b)] for a, b in re.findall(r"Vec2<deviceFloat>\{([-0-9.]+),\s*([-0-9.]+)\}", block)]


def main():
    text = open(SRC).read()
    cases = []
    for m in re.finditer(r'TEST_CASE\("([^"]+)"', text):
        start = m.end()
        nxt = text.find("TEST_CASE(", start)
        body = text[start:nxt if nxt >= 0 else len(text)]
        pts = vectors(body[body.index("points {"):body.index("local_coordinates {")])
        loc = vectors(body[body.index("local_coordinates {"):body.index("global_coordinates {")])
        glo = vectors(body[body.index("global_coordinates {"):body.index("local_coordinates_computed;")])
        tol = float(re.search(r"const double error = ([0-9.e-]+);", body).group(1))
        line = text.count("\n", 0, m.start()) + 1
        cases.append({"name": m.group(1), "line": line, "points": pts, "local": loc, "global": glo, "tolerance": tol})
    out = {"source": "tests/inverse_quad_mapping_test.cu", "inverse_quad_map": cases}
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    sys.exit(main())
