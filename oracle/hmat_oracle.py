"""CPU restatement of the reference's H-matrix tree rules and block product
(test infrastructure only: imported by tests/ and nothing else).

* ``cluster_leaves(points, min_size)``: the cluster tree of
  src/Hmatrix.py:153-192 (split at the midpoint of the largest bounding-box
  extent, ``<= gamma`` to the left, stop at <= min_size points or an empty
  side), returned as an explicit node table;
* ``block_leaves(points, min_size, mini_leaf, eta)``: the block tree of
  src/Hmatrix.py:2077-2142 with the admissibility test :66-113, leaves in the
  depth-first order of ``collect_blocks`` (:1618-1624), each with the
  reference's ACA/dense decision of ``ACA_worker`` (:1296, both clusters
  larger than mini_leaf -> low rank);
* ``blocks_mvm(x, leaves, type)``: the reference's y[rows] += U (V x[cols])
  or Mf x[cols] accumulation (src/Hmatrix.py:1773-1866) in FP64;
* ``dense_expansion(leaves, n)``: the six n x n matrices an H-matrix
  represents.

Pins (tests/test_hmat_cpu.py): hand-worked splits and admissibility on tiny
point sets (ties at the midpoint, degenerate axes), the partition invariant
(every matrix entry in exactly one leaf), admissibility of every low-rank
leaf, and exact agreement of the leaf list with the product's tree.
"""
import numpy as np


def _box(points, idx):
    p = points[idx]
    return p.min(axis=0), p.max(axis=0)


def cluster_leaves(points, min_size=16):
    """Node table of the cluster tree: list of dicts with 'idx' (ascending
    element ids), 'lo', 'hi', 'kids' (two node numbers or None)."""
    points = np.asarray(points, dtype=np.float64)
    nodes = [{"idx": np.arange(len(points)), "kids": None}]
    todo = [0]
    while todo:
        k = todo.pop()
        idx = nodes[k]["idx"]
        lo, hi = _box(points, idx)
        nodes[k]["lo"], nodes[k]["hi"] = lo, hi
        if idx.size <= min_size:
            continue
        ext = hi - lo
        ax = 0
        for d in range(1, ext.size):   # first axis of the largest extent
            if ext[d] > ext[ax]:
                ax = d
        gamma = (lo[ax] + hi[ax]) / 2
        left = np.array([i for i in idx if points[i, ax] <= gamma], dtype=np.int64)
        right = np.array([i for i in idx if not points[i, ax] <= gamma], dtype=np.int64)
        if left.size == 0 or right.size == 0:
            continue
        nodes.append({"idx": left, "kids": None})
        nodes.append({"idx": right, "kids": None})
        nodes[k]["kids"] = (len(nodes) - 2, len(nodes) - 1)
        todo.append(len(nodes) - 2)
        todo.append(len(nodes) - 1)
    return nodes


def admissible(lo_t, hi_t, lo_s, hi_s, eta=2.0):
    """min(diam) <= eta * dist of two boxes (src/Hmatrix.py:66-113)."""
    d_t = np.sqrt(np.sum((hi_t - lo_t) ** 2))
    d_s = np.sqrt(np.sum((hi_s - lo_s) ** 2))
    gap = 0.0
    for a_t, b_t, a_s, b_s in zip(lo_t, hi_t, lo_s, hi_s):
        g = max(0.0, a_s - b_t, a_t - b_s)   # at most one of the two is positive
        gap += g * g
    return min(d_t, d_s) <= eta * np.sqrt(gap)


def block_leaves(points, min_size=16, mini_leaf=32, eta=2.0):
    """Leaves of the block tree in depth-first order: list of
    (row ids, column ids, lowrank flag)."""
    nodes = cluster_leaves(points, min_size)
    out = []
    stack = [(0, 0)]
    while stack:
        r, c = stack.pop()
        R, C = nodes[r], nodes[c]
        leaf = (admissible(R["lo"], R["hi"], C["lo"], C["hi"], eta) or R["idx"].size <= mini_leaf
                or C["idx"].size <= mini_leaf or R["kids"] is None or C["kids"] is None)
        if leaf:
            lowrank = R["idx"].size > mini_leaf and C["idx"].size > mini_leaf
            out.append((R["idx"], C["idx"], lowrank))
            continue
        rl, rr = R["kids"]
        cl, cr = C["kids"]
        # children [LL, LR, RL, RR], visited in that order
        for pair in ((rr, cr), (rr, cl), (rl, cr), (rl, cl)):
            stack.append(pair)
    return out


NAMES = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")


def blocks_mvm(x, leaves, type):
    """Reference block product (src/Hmatrix.py:1773-1866).  ``leaves``: list
    of (rows, cols, dense dict or None, factors dict or None) where dense
    maps name -> m x n and factors maps name -> (U, V)."""
    y = np.zeros(len(x))
    for rows, cols, dense, fac in leaves:
        xs = x[cols]
        if fac is not None:
            U, V = fac[type]
            part = U @ (V @ xs) if U.shape[1] else np.zeros(len(rows))
        else:
            part = dense[type] @ xs
        y[rows] += part
    return y


def dense_expansion(leaves, n):
    """The six n x n FP64 matrices represented by the leaves."""
    mats = [np.zeros((n, n)) for _ in NAMES]
    for rows, cols, dense, fac in leaves:
        for k, name in enumerate(NAMES):
            if fac is not None:
                U, V = fac[name]
                mats[k][np.ix_(rows, cols)] = U @ V if U.shape[1] else 0.0
            else:
                mats[k][np.ix_(rows, cols)] = dense[name]
    return mats
