"""Mesh tools around the drop-in path: SU2 writing and element renumbering.

The reference renumbers SU2 meshes offline so that consecutive elements are close in space
(extras/mesh_renumberer/mesh_renumberer.py:88-155, hilbert.py:4-40): "circular" sorts elements by
the angle of their centre around the origin, "hilbert-circular" maps the centres from the disk
that holds them onto the square [-1, 1]^2 and sorts them by their Hilbert index on an n x n grid
(n the smallest power of two above the element count). The element order is the storage order
of the solver, so it sets memory locality on the GPU.

    python -m ndg_cuda_amd.mesh_tools renumber -i in.su2 -o out.su2 [-a circular|hilbert-circular]
"""
import argparse
import sys

import numpy as np

from .sem import BC_INFLOW, BC_OUTFLOW, BC_SYMMETRY, BC_WALL, MeshDescription, read_su2

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


def write_su2(desc: MeshDescription, path) -> None:
    """Writes a description without periodic bars as an SU2 file (one marker per boundary kind,
    in the order the kinds first appear); read back by read_su2 / Mesh2D_t(path, ...)."""
    bar_type = np.asarray(desc.bar_type)
    if any(int(t) not in SU2_MARKERS for t in bar_type):
        raise ValueError("write_su2: SU2 files hold wall, symmetry, inlet and farfield markers only")
    lines = ["NDIME= 2", f"NELEM= {len(desc.quads)}"]
    lines += ["9 " + " ".join(str(int(v)) for v in q) + f" {i}" for i, q in enumerate(np.asarray(desc.quads))]
    lines.append(f"NPOIN= {len(desc.xy)}")
    lines += [f"{x!r} {y!r} {i}" for i, (x, y) in enumerate(np.asarray(desc.xy).tolist())]
    kinds = list(dict.fromkeys(int(t) for t in bar_type))
    lines.append(f"NMARK= {len(kinds)}")
    for k in kinds:
        sel = np.nonzero(bar_type == k)[0]
        lines += [f"MARKER_TAG= {SU2_MARKERS[k]}", f"MARKER_ELEMS= {len(sel)}"]
        lines += [f"3 {int(desc.bars[b][0])} {int(desc.bars[b][1])}" for b in sel]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m ndg_cuda_amd.mesh_tools")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("renumber", help="re-numbers the elements of an SU2 mesh")
    r.add_argument("-i", "--input", required=True, help="SU2 mesh to re-number")
    r.add_argument("-o", "--output", required=True, help="path of the re-numbered mesh")
    r.add_argument("-a", "--algorithm", default="circular", choices=["circular", "hilbert-circular"])
    a = ap.parse_args(argv)
    write_su2(renumber(read_su2(a.input), a.algorithm), a.output)
    return 0


if __name__ == "__main__":
    sys.exit(main())
