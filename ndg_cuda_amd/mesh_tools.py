"""Mesh tools around the drop-in path: SU2 writing and element renumbering.

The reference renumbers SU2 meshes offline so that consecutive elements are close in space
(extras/mesh_renumberer/mesh_renumberer.py:88-155, hilbert.py:4-40): "circular" sorts elements by
the angle of their centre around the origin, "hilbert-circular" maps the centres from the disk
that holds them onto the square [-1, 1]^2 and sorts them by their Hilbert index on an n x n grid
(n the smallest power of two above the element count). The element order is the storage order
of the solver, so it sets memory locality on the GPU.

    python -m ndg_cuda_amd.mesh_tools renumber -i in.su2 -o out.su2 [-a circular|hilbert-circular]
                                             [-p TAG_A:TAG_B ...]

SU2 files carry no periodic pairing; write_su2 names each pair of periodic boundaries
periodic_<k>_a / periodic_<k>_b and read_su2(path, periodic=[(a, b), ...]) joins them again.
"""
import argparse
import sys

import numpy as np

from .sem import BC_INFLOW, BC_OUTFLOW, BC_PERIODIC, BC_SYMMETRY, BC_WALL, MeshDescription, read_su2

SU2_MARKERS = {BC_WALL: "wall", BC_SYMMETRY: "symmetry", BC_INFLOW: "inlet", BC_OUTFLOW: "farfield"}


def element_centers(desc: MeshDescription) -> np.ndarray:
    """Mean of the four corner nodes of every element (mesh_renumberer.py:88-92)."""
    return np.asarray(desc.xy)[np.asarray(desc.quads)].mean(axis=1)


def circular_order(centers: np.ndarray) -> np.ndarray:
    """Elements by the angle atan2(y, x) of their centre (mesh_renumberer.py:94-98); stable."""
    return np.argsort(np.arctan2(centers[:, 1], centers[:, 0]), kind="stable")


def disk_to_square(centers: np.ndarray) -> np.ndarray:
    """The centres scaled into the unit disk and mapped onto [-1, 1]^2 by the elliptical grid
    mapping's inverse (mesh_renumberer.py:100-107)."""
    radius = np.max(np.hypot(centers[:, 0], centers[:, 1]))
    x, y = centers[:, 0] / radius, centers[:, 1] / radius
    r2 = 2 * np.sqrt(2)
    u = np.sqrt(np.maximum(2 + x * x - y * y + x * r2, 0)) / 2 - np.sqrt(np.maximum(2 + x * x - y * y - x * r2, 0)) / 2
    v = np.sqrt(np.maximum(2 - x * x + y * y + y * r2, 0)) / 2 - np.sqrt(np.maximum(2 - x * x + y * y - y * r2, 0)) / 2
    return np.column_stack([u, v])


def hilbert_xy2d(n: int, x: int, y: int) -> int:
    """Distance along the Hilbert curve of cell (x, y) of an n x n grid, n a power of two
    (hilbert.py:4-16, the inverse of the square-mesh generator's d2xy)."""
    d = 0
    s = n // 2
    while s > 0:
        rx = 1 if x & s else 0
        ry = 1 if y & s else 0
        d += s * s * ((3 * rx) ^ ry)
        if ry == 0:  # rotate the quadrant (hilbert.py:32-40)
            if rx == 1:
                x, y = n - 1 - x, n - 1 - y
            x, y = y, x
        s //= 2
    return d


def hilbert_circular_order(centers: np.ndarray) -> np.ndarray:
    """Elements by the Hilbert index of their disk-to-square mapped centre (mesh_renumberer.py:
    109-122); stable."""
    sq = disk_to_square(centers)
    n = 2
    while n <= centers.shape[0]:
        n *= 2
    # a coordinate of n (a centre mapped onto the square's upper edge) has the bits of 0, as in
    # the reference's bit tests
    cells = np.rint((sq + 1) * n / 2).astype(np.int64) % n
    d = np.array([hilbert_xy2d(n, int(cx), int(cy)) for cx, cy in cells], dtype=np.int64)
    return np.argsort(d, kind="stable")


def renumber(desc: MeshDescription, algorithm: str = "circular") -> MeshDescription:
    """The description with its elements (and per-element orders) in the chosen order; nodes
    and boundary elements are unchanged."""
    centers = element_centers(desc)
    if algorithm == "circular":
        order = circular_order(centers)
    elif algorithm == "hilbert-circular":
        order = hilbert_circular_order(centers)
    else:
        raise ValueError(f"unknown algorithm '{algorithm}', only 'circular' and 'hilbert-circular'")
    quad_N = None if desc.quad_N is None else np.asarray(desc.quad_N)[order]
    return MeshDescription(np.asarray(desc.xy).copy(), np.asarray(desc.quads)[order].copy(),
                           np.asarray(desc.bars).copy(), np.asarray(desc.bar_type).copy(),
                           np.asarray(desc.bar_donor).copy(), quad_N)


def periodic_groups(desc: MeshDescription):
    """The periodic bars of a description grouped as marker pairs: bars whose donor lies at the
    same translation t form marker a, their donors (at -t) marker b. Returns a list of
    (bars of a, bars of b), each in bar order."""
    bar_type = np.asarray(desc.bar_type)
    per = np.nonzero(bar_type == BC_PERIODIC)[0]
    if per.size == 0:
        return []
    xy, bars = np.asarray(desc.xy), np.asarray(desc.bars)
    donor = np.asarray(desc.bar_donor)
    mid = (xy[bars[:, 0]] + xy[bars[:, 1]]) / 2
    scale = float(np.max(np.abs(xy))) or 1.0
    groups = {}
    for b in per:
        d = int(donor[b])
        if d < 0 or bar_type[d] != BC_PERIODIC or int(donor[d]) != b:
            raise ValueError("write_su2: periodic bars must come in mutual donor pairs")
        key = tuple(np.round((mid[d] - mid[b]) / scale, 9))
        groups.setdefault(key, []).append(int(b))
    pairs, used = [], set()
    for key, members in groups.items():
        neg = tuple(-k + 0.0 for k in key)
        if key in used:
            continue
        if neg not in groups:
            raise ValueError("write_su2: a periodic group has no opposite group")
        used.update({key, neg})
        pairs.append((members, groups[neg]))
    return pairs


def write_su2(desc: MeshDescription, path):
    """Writes a description as an SU2 file (one marker per boundary kind, in the order the kinds
    first appear), read back by read_su2 / Mesh2D_t(path, ...). Periodic bars become marker pairs
    "periodic_<k>_a" / "periodic_<k>_b" (one pair per translation); the list of those tag pairs is
    returned, to be given to read_su2(path, periodic=...)."""
    bar_type = np.asarray(desc.bar_type)
    if any(int(t) not in SU2_MARKERS and int(t) != BC_PERIODIC for t in bar_type):
        raise ValueError("write_su2: SU2 files hold wall, symmetry, inlet, farfield and periodic markers only")
    groups = periodic_groups(desc)
    lines = ["NDIME= 2", f"NELEM= {len(desc.quads)}"]
    lines += ["9 " + " ".join(str(int(v)) for v in q) + f" {i}" for i, q in enumerate(np.asarray(desc.quads))]
    lines.append(f"NPOIN= {len(desc.xy)}")
    lines += [f"{x!r} {y!r} {i}" for i, (x, y) in enumerate(np.asarray(desc.xy).tolist())]
    kinds = list(dict.fromkeys(int(t) for t in bar_type if int(t) != BC_PERIODIC))
    lines.append(f"NMARK= {len(kinds) + 2 * len(groups)}")
    bars = np.asarray(desc.bars)

    def marker(tag, sel):
        lines.extend([f"MARKER_TAG= {tag}", f"MARKER_ELEMS= {len(sel)}"])
        lines.extend(f"3 {int(bars[b][0])} {int(bars[b][1])}" for b in sel)

    for k in kinds:
        marker(SU2_MARKERS[k], np.nonzero(bar_type == k)[0])
    tags = []
    for i, (a, b) in enumerate(groups):
        marker(f"periodic_{i}_a", a)
        marker(f"periodic_{i}_b", b)
        tags.append((f"periodic_{i}_a", f"periodic_{i}_b"))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return tags


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m ndg_cuda_amd.mesh_tools")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("renumber", help="re-numbers the elements of an SU2 mesh")
    r.add_argument("-i", "--input", required=True, help="SU2 mesh to re-number")
    r.add_argument("-o", "--output", required=True, help="path of the re-numbered mesh")
    r.add_argument("-a", "--algorithm", default="circular", choices=["circular", "hilbert-circular"])
    r.add_argument("-p", "--periodic", action="append", default=[], metavar="A:B",
                   help="marker tags A and B joined by a translation (repeatable)")
    a = ap.parse_args(argv)
    pairs = []
    for spec in a.periodic:
        if spec.count(":") != 1:
            ap.error(f"--periodic expects A:B, got '{spec}'")
        pairs.append(tuple(spec.split(":")))
    tags = write_su2(renumber(read_su2(a.input, periodic=pairs), a.algorithm), a.output)
    for ta, tb in tags:
        print(f"periodic markers in {a.output}: {ta}:{tb}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
