"""Tree-quality figures of a session's resident BVH (`Session.bvh()`).

The surface-area heuristic cost of a 4-wide tree, relative to a ray that
hits its root box: the sum over its nodes of area(node) / area(root) (one
node visit each, four box tests) plus the sum over its leaf slots of
area(slot) / area(root) x primitives in the slot (one primitive test each,
an instance entry in the top level). The host build computes the same sum
while it collapses (bvh.cpp `sahCost`); here it is taken from the node
bytes the traversal kernels read, so every builder is costed the same way.
"""

import numpy as np

LEAF = 0x80000000
EMPTY = 0xFFFFFFFF


def node_fields(raw):
    """(lo (N, 4, 3), hi (N, 4, 3), ref (N, 4) uint32, count (N, 4) int32)
    of raw 128-byte nodes: rows lo.x, hi.x, lo.y, hi.y, lo.z, hi.z, refs,
    counts, four children per row."""
    f = np.frombuffer(raw, np.float32).reshape(-1, 8, 4)
    lo = np.stack([f[:, 0], f[:, 2], f[:, 4]], axis=-1)
    hi = np.stack([f[:, 1], f[:, 3], f[:, 5]], axis=-1)
    ref = f[:, 6].view(np.uint32)
    cnt = f[:, 7].view(np.int32)
    return lo, hi, ref, cnt


def _area(lo, hi):
    d = (hi - lo).astype(np.float64)
    return d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2] + d[..., 2] * d[..., 0]


def sah_cost(fields, root):
    """SAH cost of the 4-wide tree under node index `root` (leaf slots are
    not descended: a top-level leaf counts as one instance entry)."""
    lo, hi, ref, _ = fields
    total = 0.0
    root_area = None
    stack = [int(root)]
    while stack:
        n = stack.pop()
        valid = ref[n] != EMPTY
        a = float(_area(lo[n][valid].min(axis=0), hi[n][valid].max(axis=0)))
        if root_area is None:
            root_area = a if a > 0.0 else 1.0
        total += a / root_area
        for k in np.nonzero(valid)[0]:
            r = int(ref[n, k])
            if r & LEAF:
                total += float(_area(lo[n, k], hi[n, k])) / root_area * (((r >> 27) & 15) + 1)
            else:
                stack.append(r)
    return total


def summary(bvh):
    """{"tlas_sah", "blas_sah" (per distinct mesh root), "nodes", "tlas_nodes"}
    of a `Session.bvh()` record."""
    fields = node_fields(bvh["nodes"])
    roots = sorted({int(r) for r in bvh["blas_roots"] if r >= 0})
    tlas_root = bvh["tlas_root"]
    out = {"nodes": bvh["num_nodes"], "tlas_nodes": bvh["tlas_nodes"],
           "blas_sah": {r: sah_cost(fields, r) for r in roots}}
    # a top level that is a single leaf reference has no node to cost
    out["tlas_sah"] = None if tlas_root & LEAF else sah_cost(fields, tlas_root)
    return out
