"""SAH cost of a 4-wide node array (sundog_amd.bvhinfo) on a hand-built
two-level tree whose cost is worked out by hand."""

import numpy as np

from sundog_amd import bvhinfo


def _node(slots):
    """128-byte node from up to 4 (lo, hi, ref, count) slots."""
    rows = np.full((8, 4), 3e38, np.float32)
    refs = np.full(4, bvhinfo.EMPTY, np.uint32)
    cnts = np.full(4, -1, np.int32)
    for k, (lo, hi, ref, cnt) in enumerate(slots):
        for a in range(3):
            rows[2 * a, k] = lo[a]
            rows[2 * a + 1, k] = hi[a]
        refs[k] = ref
        cnts[k] = cnt
    rows[6] = refs.view(np.float32)
    rows[7] = cnts.view(np.float32)
    return rows.tobytes()


def _leaf(first, count):
    return bvhinfo.LEAF | ((count - 1) << 27) | first


def test_sah_cost_of_a_hand_built_tree():
    # root: an inner child over [0,2]x[0,1]x[0,1] and a 2-primitive leaf
    # over [2,4]x[0,1]x[0,1]; the child: leaves of 1 and 3 primitives over
    # the unit cubes at x in [0,1] and [1,2]. Half areas: root box 9, child
    # 5, leaves 5, 3, 3. Cost = (9 + 5 + 2*5 + 1*3 + 3*3) / 9 = 4.
    root = _node([((0, 0, 0), (2, 1, 1), 1, 0), ((2, 0, 0), (4, 1, 1), _leaf(0, 2), 2)])
    child = _node([((0, 0, 0), (1, 1, 1), _leaf(2, 1), 1), ((1, 0, 0), (2, 1, 1), _leaf(3, 3), 3)])
    fields = bvhinfo.node_fields(root + child)
    assert abs(bvhinfo.sah_cost(fields, 0) - 4.0) < 1e-12
    # the child alone, costed relative to its own box: (5 + 3 + 9) / 5
    assert abs(bvhinfo.sah_cost(fields, 1) - 17.0 / 5.0) < 1e-12


def test_summary_separates_levels():
    top = _node([((0, 0, 0), (1, 1, 1), _leaf(0, 1), 1), ((1, 0, 0), (2, 1, 1), _leaf(1, 1), 1)])
    mesh = _node([((0, 0, 0), (1, 1, 1), _leaf(0, 4), 4)])
    rec = {"nodes": top + mesh, "num_nodes": 2, "tlas_nodes": 1, "tlas_root": 0,
           "blas_roots": np.array([1, 1, -1], np.int32)}
    s = bvhinfo.summary(rec)
    # top: (5 + 3 + 3) / 5; mesh: (3 + 4 * 3) / 3
    assert abs(s["tlas_sah"] - 11.0 / 5.0) < 1e-12
    assert list(s["blas_sah"]) == [1] and abs(s["blas_sah"][1] - 5.0) < 1e-12
