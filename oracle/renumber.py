"""Edge ordering of a compiled program at world size 1.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference's compiler keeps, for every edge space bound to a Reducer, the
order of the Reducer's index stable-sorted (the reorder plan keeps the
Reducer edge, passes/sparse_encoding/reorder_plan.py:230-231; the permutation
is zipsort_using_order with a stable sort, passes/utils.py:301-336, applied
in passes/sparse_encoding/sparse_encoding.py:298-303). At world size 1 no
reordering Selector is added (sparse_encoding.py:418-424), and vertex/cell
spaces keep their order. So edge k of the compiled program is the
k-th entry of argsort(reducer_idx, stable), and a row's edges keep the user's
relative order -- the per-row summation order of the Reducer.
"""
import numpy as np


def edge_order(reducer_idx, n_rows=None):
    """Permutation p with reducer_idx[p] sorted and ties in input order,
    written as a counting sort (bucket per row, rows in increasing order)."""
    idx = np.asarray(reducer_idx, dtype=np.int64)
    n = int(idx.max()) + 1 if n_rows is None else int(n_rows)
    buckets = [[] for _ in range(n)]
    for e, r in enumerate(idx.tolist()):
        buckets[r].append(e)
    return np.array([e for b in buckets for e in b], dtype=np.int64)


def row_edge_lists(reducer_idx, order, n_rows):
    """For every row, the user's edge ids of that row in compiled order
    (order[k] = user edge id at compiled position k)."""
    idx = np.asarray(reducer_idx, dtype=np.int64)
    out = [[] for _ in range(n_rows)]
    for e in np.asarray(order).tolist():
        out[idx[e]].append(e)
    return out
