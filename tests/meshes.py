"""Mesh descriptions with hanging (2:1 non-conforming) faces, for the tests.

split_elements replaces each listed element by its four children (edge midpoints and the centre
as new nodes, each child keeping the parent's local orientation, children in the parent's place);
a side of a split element whose neighbour is not split becomes a hanging side of that neighbour.
Only elements without boundary bars may be split (the bars stay as they are).
"""
import numpy as np

import ndg_cuda_amd as nd


def split_elements(desc, elements):
    elements = set(int(e) for e in elements)
    xy = [tuple(p) for p in np.asarray(desc.xy)]
    quads = np.asarray(desc.quads)
    bar_edges = {frozenset(map(int, b)) for b in np.asarray(desc.bars)}
    mid = {}

    def node(p):
        xy.append(p)
        return len(xy) - 1

    def midpoint(a, b):
        key = (min(a, b), max(a, b))
        if key not in mid:
            pa, pb = xy[a], xy[b]
            mid[key] = node(((pa[0] + pb[0]) / 2, (pa[1] + pb[1]) / 2))
        return mid[key]

    out, orders = [], []
    qn = None if desc.quad_N is None else np.asarray(desc.quad_N)
    for e, q in enumerate(quads):
        q = [int(v) for v in q]
        if e not in elements:
            out.append(q)
            orders.append(None if qn is None else int(qn[e]))
            continue
        for s in range(4):
            assert frozenset((q[s], q[(s + 1) % 4])) not in bar_edges, "split elements must be interior"
        m01, m12, m23, m30 = (midpoint(q[s], q[(s + 1) % 4]) for s in range(4))
        c = node(tuple(np.mean([xy[v] for v in q], axis=0)))
        for child in ([q[0], m01, c, m30], [m01, q[1], m12, c], [c, m12, q[2], m23], [m30, c, m23, q[3]]):
            out.append(child)
            orders.append(None if qn is None else int(qn[e]))
    return nd.MeshDescription(np.array(xy, dtype=np.float64), np.array(out, dtype=np.int64),
                              np.asarray(desc.bars).copy(), np.asarray(desc.bar_type).copy(),
                              np.asarray(desc.bar_donor).copy(),
                              None if qn is None else np.array(orders, dtype=np.int32))


def interior_elements(desc):
    """Elements with no node on the boundary bars."""
    bar_nodes = set(np.asarray(desc.bars).ravel().tolist())
    return [e for e, q in enumerate(np.asarray(desc.quads)) if not bar_nodes.intersection(q.tolist())]


SU2_TAGS = {nd.BC_WALL: "wall", nd.BC_SYMMETRY: "symmetry", nd.BC_INFLOW: "inlet", nd.BC_OUTFLOW: "farfield"}


def write_su2(desc, path, tags=None):
    """Writes a description without periodic bars as an SU2 file (one marker per boundary kind,
    in the order the kinds first appear), the format the reference's Mesh2D_t::read_su2 reads."""
    tags = tags or SU2_TAGS
    bar_type = np.asarray(desc.bar_type)
    lines = ["NDIME= 2", f"NPOIN= {len(desc.xy)}"]
    lines += [f"{x!r} {y!r} {i}" for i, (x, y) in enumerate(np.asarray(desc.xy).tolist())]
    lines.append(f"NELEM= {len(desc.quads)}")
    lines += ["9 " + " ".join(str(int(v)) for v in q) + f" {i}" for i, q in enumerate(np.asarray(desc.quads))]
    kinds = list(dict.fromkeys(int(t) for t in bar_type))
    lines.append(f"NMARK= {len(kinds)}")
    for k in kinds:
        sel = np.nonzero(bar_type == k)[0]
        lines += [f"MARKER_TAG= {tags[k]}", f"MARKER_ELEMS= {len(sel)}"]
        lines += [f"3 {int(desc.bars[b][0])} {int(desc.bars[b][1])}" for b in sel]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
