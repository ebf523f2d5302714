"""Hierarchical-matrix traction Green's functions (the reference's src/Hmatrix.py
interface, backed by libpq3d.so).

Layers, each mirroring the reference's names and argument meaning:

* cluster tree -- ``build_block_tree`` (Hmatrix.py:153-192): split a cluster at
  the midpoint of its bounding box's largest extent (``<= gamma`` left,
  ``> gamma`` right) until it has at most ``min_size`` points or a side would
  be empty;
* block tree -- ``create_recursive_blocks`` (:2077-2142): a (row, column)
  cluster pair is a leaf if it is admissible (``is_admissible``, :98-113,
  min(diam) <= eta * dist of the bounding boxes) or either cluster has at most
  ``mini_leaf`` points (or cannot be split); otherwise four children
  [LL, LR, RL, RR];
* ``createHmatrix`` (:2147-2196) builds both cluster trees from the element
  centroids and the block tree, and ``BlockTree.collect_blocks`` (:1618-1624)
  lists the leaves in depth-first order;
* ``BlockTree.ACA_worker`` (:1292-1344): a leaf with both clusters larger than
  ``mini_leaf`` gets low-rank factors U (m x k), V (k x n) per matrix in
  ``ACA_dictS``/``ACA_dictD`` (keys ``U_ACA_A1s`` ... ``V_ACA_Bd``) and
  ``judaca = True``; any other leaf gets its six dense sub-blocks ``Mf_*``;
* ``BlockTree.blocks_process_MVM(x, blocks, type)`` (:1773-1866): y = A_type x.

Where the work runs: the matrix entries of every ACA pivot row/column and of
every dense leaf are evaluated on the GPU (one batched launch per ACA sweep
over all blocks, ``pq3d_gf_segments``); the ACA pivoting, residual updates
and recompression run on the GPU too (``pq3d_aca_run_device``, the default;
``pq3d_aca_run`` keeps the pivoting on host threads); ``DeviceHmatrix`` packs
the factors and dense leaves into a device-resident, cluster-ordered layout
and runs the product with HIP kernels (``pq3d_hmat_*``), which is also what
the stepping path uses.

ACA stopping rule: the reference stops on an absolute step size of 0.01 (in
Pa/m of slip, :1301-1321) or a pivot below 1e-4; here the rule is relative,
||u_k|| ||v_k|| <= eps ||A_k||_F with a reference row and column that must
also be resolved to eps (eps = 1e-4 by default), so the accuracy does not
depend on the mesh's units or element size.  The pivots are not the
reference's (it draws its start rows from an unseeded RNG, :740), so factors
match the reference only through the accuracy bound.
"""
import ctypes
import os
import math
import time

import numpy as np

from . import _lib, readmsh
from ._lib import check, ptr

MATRIX_NAMES = ("A1s", "A2s", "Bs", "A1d", "A2d", "Bd")
_BLOCK_LISTS = frozenset([p + s for p in ("U_", "S_", "Vt_") for s in ("1s", "2s", "Bs", "1d", "2d", "Bd")]
                         + ["Mf_" + n for n in MATRIX_NAMES])


# ----------------------------------------------------------------- geometry
def bounding_box(cluster, points):
    """Lower and upper corner of the bounding box of points[cluster]."""
    p = points[cluster]
    return p.min(axis=0), p.max(axis=0)


def diameter(a, b):
    """Diagonal length of the box [a, b].  Python floats, summed over the
    axes in order: the same roundings as NumPy's sum of three squares, without
    its per-call overhead (the block tree evaluates this ~10^5 times)."""
    s = 0.0
    for x, y in zip(np.asarray(a, dtype=np.float64).tolist(), np.asarray(b, dtype=np.float64).tolist()):
        d = y - x
        s += d * d
    return math.sqrt(s)


def distance(a_tau, b_tau, a_sigma, b_sigma):
    """Euclidean distance between the boxes [a_tau, b_tau] and [a_sigma, b_sigma]."""
    return _box_distance(*(np.asarray(v, dtype=np.float64).tolist() for v in (a_tau, b_tau, a_sigma, b_sigma)))


def _box_distance(a_tau, b_tau, a_sigma, b_sigma):
    """``distance`` on lists of floats: per axis max(0, a_s - b_t)^2 +
    max(0, a_t - b_s)^2, summed over the axes in order."""
    s = 0.0
    for at, bt, as_, bs in zip(a_tau, b_tau, a_sigma, b_sigma):
        g1 = as_ - bt
        g2 = at - bs
        g1 = g1 if g1 > 0.0 else 0.0
        g2 = g2 if g2 > 0.0 else 0.0
        s += g1 * g1 + g2 * g2
    return math.sqrt(s)


def _admissible_boxes(lo_t, hi_t, lo_s, hi_s, eta):
    return min(diameter(lo_t, hi_t), diameter(lo_s, hi_s)) <= eta * distance(lo_t, hi_t, lo_s, hi_s)


def _admissible_nodes(t, s, eta):
    """``_admissible_boxes`` for two cluster-tree nodes, from their cached
    corners and diameters."""
    lo_t, hi_t, d_t = t.geometry()
    lo_s, hi_s, d_s = s.geometry()
    return min(d_t, d_s) <= eta * _box_distance(lo_t, hi_t, lo_s, hi_s)


def is_admissible(cluster_tau, cluster_sigma, points, eta=2.0):
    """True when min(diam(tau), diam(sigma)) <= eta * dist(tau, sigma) for the
    bounding boxes of the two clusters (Hmatrix.py:98-113)."""
    lo_t, hi_t = bounding_box(cluster_tau, points)
    lo_s, hi_s = bounding_box(cluster_sigma, points)
    return _admissible_boxes(lo_t, hi_t, lo_s, hi_s, eta)


# ----------------------------------------------------------------- cluster tree
class TreeNode:
    """Cluster-tree node: ``indices`` (ascending element ids), split axis and
    threshold, children; plus its bounding box and ``start``, the offset of
    its range in the depth-first leaf order (every cluster is a contiguous
    range of that ordering)."""

    def __init__(self, indices, split_dim=None, split_value=None, level=0):
        self.indices = indices
        self.split_dim = split_dim
        self.split_value = split_value
        self.level = level
        self.left = None
        self.right = None
        self.lo = None
        self.hi = None
        self.start = 0
        self._geometry = None

    def geometry(self):
        """(lo, hi, diameter) of the bounding box, corners as lists of floats;
        computed once per node."""
        g = self._geometry
        if g is None:
            lo, hi = self.lo.tolist(), self.hi.tolist()
            g = self._geometry = (lo, hi, diameter(lo, hi))
        return g


def build_block_tree(cluster, points, min_size=16, depth=0):
    """Binary cluster tree over ``cluster`` (indices into ``points``), split
    rules of Hmatrix.py:153-192.  Returns the root ``TreeNode``."""
    cluster = np.asarray(cluster, dtype=np.int64)
    points = np.asarray(points, dtype=np.float64)

    def make(idx, level):
        p = points[idx]
        lo, hi = p.min(axis=0), p.max(axis=0)
        if len(idx) <= min_size:
            node = TreeNode(idx, level=level)
        else:
            axis = int(np.argmax(hi - lo))
            gamma = (lo[axis] + hi[axis]) / 2
            mask = p[:, axis] <= gamma
            left, right = idx[mask], idx[~mask]
            if len(left) == 0 or len(right) == 0:
                node = TreeNode(idx, level=level)
            else:
                node = TreeNode(idx, split_dim=axis, split_value=gamma, level=level)
                node.left = make(left, level + 1)
                node.right = make(right, level + 1)
        node.lo, node.hi = lo, hi
        return node

    root = make(cluster, depth)
    _assign_starts(root)
    return root


def _assign_starts(node, pos=0):
    """Set ``start`` of every node (offset in the depth-first leaf order);
    returns the offset after the subtree."""
    node.start = pos
    if node.left is None:
        return pos + len(node.indices)
    return _assign_starts(node.right, _assign_starts(node.left, pos))


def leaf_order(root):
    """Element ids in depth-first leaf order (the cluster ordering)."""
    out = []
    stack = [root]
    while stack:
        node = stack.pop()
        if node.left is None:
            out.append(node.indices)
        else:
            stack.append(node.right)
            stack.append(node.left)
    return np.concatenate(out)


# ----------------------------------------------------------------- blocks
class Block:
    """Leaf or interior block of the block tree (attributes of Hmatrix.py:204-256)."""

    def __init__(self, row_cluster, col_cluster, row_index, col_index, children=None, level=0):
        self.row_cluster = row_cluster
        self.col_cluster = col_cluster
        self.row_index = row_index
        self.col_index = col_index
        self.ACA_dictS = {}
        self.ACA_dictD = {}
        self.children = children if children is not None else []
        self.level = level
        self.judproc = False
        self.judsvd = False
        self.judaca = False
        self.row_node = None
        self.col_node = None

    def __getattr__(self, name):
        # The per-matrix lists U_*, S_*, Vt_* and Mf_* start empty; each is
        # made on first access, so the ~4*10^4 blocks of a large tree do not
        # allocate 24 lists apiece up front.
        if name in _BLOCK_LISTS:
            v = []
            setattr(self, name, v)
            return v
        raise AttributeError(name)

    def setdata(self, data1):
        self.data = data1

    def is_leaf(self):
        return len(self.children) == 0

    def factors(self, name):
        """(U, V) of a low-rank block for matrix ``name`` (empty arrays for an
        identically zero matrix)."""
        d = self.ACA_dictS if name.endswith("s") else self.ACA_dictD
        return d["U_ACA_" + name], d["V_ACA_" + name]


def create_recursive_blocks(row_cluster, col_cluster, row_index, col_index, points, plotHmatrix=False,
                            mini_leaf=16, depth=0, eta=2.0):
    """Block tree over a pair of cluster-tree nodes (Hmatrix.py:2077-2142)."""
    def leaf():
        b = Block(row_cluster.indices, col_cluster.indices, row_index, col_index, level=depth)
        b.row_node, b.col_node = row_cluster, col_cluster
        return b

    if (_admissible_nodes(row_cluster, col_cluster, eta)
            or len(row_cluster.indices) <= mini_leaf or len(col_cluster.indices) <= mini_leaf):
        return leaf()
    if (row_cluster.left is None or row_cluster.right is None or col_cluster.left is None
            or col_cluster.right is None):
        return leaf()
    half = int(len(row_index) / 2)
    rl, rr = row_index[:half], row_index[half:]
    cl, cr = col_index[:half], col_index[half:]
    children = [
        create_recursive_blocks(row_cluster.left, col_cluster.left, rl, cl, points, plotHmatrix, mini_leaf,
                                depth + 1, eta),
        create_recursive_blocks(row_cluster.left, col_cluster.right, rl, cr, points, plotHmatrix, mini_leaf,
                                depth + 1, eta),
        create_recursive_blocks(row_cluster.right, col_cluster.left, rr, cl, points, plotHmatrix, mini_leaf,
                                depth + 1, eta),
        create_recursive_blocks(row_cluster.right, col_cluster.right, rr, cr, points, plotHmatrix, mini_leaf,
                                depth + 1, eta),
    ]
    b = Block(row_cluster.indices, col_cluster.indices, row_index, col_index, children=children, level=depth)
    b.row_node, b.col_node = row_cluster, col_cluster
    return b


def createHmatrix(xg, nodelst, elelst, eleVec, Para0, mini_leaf=32, min_size=16, eta=2.0):
    """Cluster trees over the element centroids and the block tree
    (Hmatrix.py:2147-2196).  ``Para0`` keys used: 'Shear modulus',
    'Lame constants', 'Half space' (as in the reference)."""
    mu, lam = float(Para0["Shear modulus"]), float(Para0["Lame constants"])
    half = Para0["Half space"]
    halfspace = half if isinstance(half, bool) else str(half).strip() == "True"
    xg = np.asarray(xg, dtype=np.float64)
    cluster = np.arange(len(xg))
    rows = build_block_tree(cluster, xg, min_size)
    cols = rows  # the reference builds the same tree twice from the same points
    root = create_recursive_blocks(rows, cols, cluster, cluster, xg, False, mini_leaf, 0, eta)
    return BlockTree(root, nodelst, elelst, eleVec, mu, lam, xg, halfspace, mini_leaf, cluster_root=rows)


# ----------------------------------------------------------------- native binding
def _declare(L):
    vp, ll, i, d = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    lp = ctypes.POINTER(ctypes.c_longlong)
    ip = ctypes.POINTER(ctypes.c_int)
    _lib._sig(L, "pq3d_gfsrc_create", [ctypes.POINTER(vp), dp, dp, dp, ll, dp, dp, ll, d, d, i])
    _lib._sig(L, "pq3d_gfsrc_destroy", [vp])
    _lib._sig(L, "pq3d_gf_segments", [vp, ip, ll, ip, ll, ll, lp, dp])
    _lib._sig(L, "pq3d_aca_run", [vp, ll, lp, ip, lp, ip, d, d, i, ctypes.c_ulonglong, lp, ctypes.POINTER(vp)])
    _lib._sig(L, "pq3d_aca_run_device", [vp, ll, lp, ip, lp, ip, d, d, i, ctypes.c_ulonglong, i, lp,
                                         ctypes.POINTER(vp)])
    _lib._sig(L, "pq3d_aca_device_info", [vp, dp])
    _lib._sig(L, "pq3d_lowrank_recompress", [ll, ll, ll, dp, dp, d, lp, dp, dp])
    _lib._sig(L, "pq3d_aca_stats", [vp, dp])
    _lib._sig(L, "pq3d_aca_phases", [vp, dp])
    _lib._sig(L, "pq3d_aca_fetch", [vp, dp, dp])
    _lib._sig(L, "pq3d_aca_free", [vp])
    _lib._sig(L, "pq3d_hmat_create", [ctypes.POINTER(vp), ll, ip, ll, lp, ip, lp, ip, ip, lp, dp, dp, dp, i])
    _lib._sig(L, "pq3d_hmat_destroy", [vp])
    _lib._sig(L, "pq3d_hmat_info", [vp, dp])
    _lib._sig(L, "pq3d_hmat_mvm", [vp, dp, dp, i, dp])
    _lib._sig(L, "pq3d_hmat_time_mvm", [vp, i, i, ctypes.POINTER(ctypes.c_float)])
    _lib._sig(L, "pq3d_dense_attach_hmatrix", [vp, vp])


_lib.register(_declare)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))


class GFSource:
    """Device-resident source and receiver geometry for batched evaluation of
    traction Green's function entries (rows, columns and sub-blocks)."""

    def __init__(self, nodelst, elelst, eleVec, xg, mu, lam, halfspace):
        nodes = np.asarray(nodelst, dtype=np.float64)
        tri = np.asarray(elelst, dtype=np.int64) - 1
        self.P1 = np.ascontiguousarray(nodes[tri[:, 0]])
        self.P2 = np.ascontiguousarray(nodes[tri[:, 1]])
        self.P3 = np.ascontiguousarray(nodes[tri[:, 2]])
        self.xg = np.ascontiguousarray(np.asarray(xg, dtype=np.float64))
        self.ev = np.ascontiguousarray(np.asarray(eleVec, dtype=np.float64))
        self.n = len(tri)
        h = ctypes.c_void_p()
        check(_lib.lib().pq3d_gfsrc_create(ctypes.byref(h), ptr(self.P1), ptr(self.P2), ptr(self.P3), self.n,
                                           ptr(self.xg), ptr(self.ev), len(self.xg), float(mu), float(lam),
                                           int(bool(halfspace))), "pq3d_gfsrc_create")
        self.h = h

    def segments(self, pairs):
        """Evaluate sub-blocks.  ``pairs``: list of (row ids, column ids).
        Returns a list of (6, m, n) float64 arrays [A1s, A2s, Bs, A1d, A2d, Bd]."""
        rows = [_i32(r) for r, _ in pairs]
        cols = [_i32(c) for _, c in pairs]
        ridx = np.concatenate(rows) if rows else np.zeros(0, np.int32)
        cidx = np.concatenate(cols) if cols else np.zeros(0, np.int32)
        seg = np.zeros((len(pairs), 4), dtype=np.int64)
        ro = co = 0
        for k, (r, c) in enumerate(zip(rows, cols)):
            seg[k] = (ro, len(r), co, len(c))
            ro += len(r)
            co += len(c)
        total = int(np.sum(seg[:, 1] * seg[:, 3]))
        out = np.empty(6 * total)
        check(_lib.lib().pq3d_gf_segments(self.h, _ip(ridx), len(ridx), _ip(cidx), len(cidx), len(pairs),
                                          _lp(seg), ptr(out)), "pq3d_gf_segments")
        res = []
        off = 0
        for k in range(len(pairs)):
            m, n = int(seg[k, 1]), int(seg[k, 3])
            # stored [comp][col][row]
            res.append(out[off:off + 6 * m * n].reshape(6, n, m).transpose(0, 2, 1))
            off += 6 * m * n
        return res

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().pq3d_gfsrc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def aca_blocks(src, blocks, eps=1e-4, recompress_eps=None, max_rank=0, seed=12345, device=False, kcap=0):
    """Adaptive cross approximation of the six matrices of every block in
    ``blocks`` (lists of row and column ids), recompressed to
    ``recompress_eps`` (default eps; 0 keeps the ACA factors).  Returns, per
    block, a list of six (U, V) pairs and the run statistics.

    device=True runs the pivot logic on the GPU (``pq3d_aca_run_device``, one
    wave per block and matrix, factors of up to ``kcap`` terms on the device,
    default and at most 64); a matrix needing more terms is redone by the host
    loop."""
    rc_eps = eps if recompress_eps is None else recompress_eps
    nb = len(blocks)
    rows = [_i32(b[0]) for b in blocks]
    cols = [_i32(b[1]) for b in blocks]
    roff = _i64(np.concatenate([[0], np.cumsum([len(r) for r in rows])]))
    coff = _i64(np.concatenate([[0], np.cumsum([len(c) for c in cols])]))
    ridx = np.concatenate(rows) if nb else np.zeros(0, np.int32)
    cidx = np.concatenate(cols) if nb else np.zeros(0, np.int32)
    ranks = np.zeros((nb, 6), dtype=np.int64)
    res = ctypes.c_void_p()
    L = _lib.lib()
    if device:
        check(L.pq3d_aca_run_device(src.h, nb, _lp(roff), _ip(ridx), _lp(coff), _ip(cidx), float(eps),
                                    float(rc_eps), int(max_rank), int(seed), int(kcap), _lp(ranks),
                                    ctypes.byref(res)), "pq3d_aca_run_device")
    else:
        check(L.pq3d_aca_run(src.h, nb, _lp(roff), _ip(ridx), _lp(coff), _ip(cidx), float(eps), float(rc_eps),
                             int(max_rank), int(seed), _lp(ranks), ctypes.byref(res)), "pq3d_aca_run")
    try:
        m = np.diff(roff)
        n = np.diff(coff)
        usz = int(np.sum(m[:, None] * ranks))
        vsz = int(np.sum(n[:, None] * ranks))
        U = np.empty(usz)
        V = np.empty(vsz)
        check(L.pq3d_aca_fetch(res, ptr(U), ptr(V)), "pq3d_aca_fetch")
        stats = np.zeros(8)
        check(L.pq3d_aca_stats(res, ptr(stats)), "pq3d_aca_stats")
        phases = np.zeros(4)
        check(L.pq3d_aca_phases(res, ptr(phases)), "pq3d_aca_phases")
        dinfo = np.zeros(4)
        check(L.pq3d_aca_device_info(res, ptr(dinfo)), "pq3d_aca_device_info")
    finally:
        L.pq3d_aca_free(res)
    out = []
    uo = vo = 0
    for b in range(nb):
        per = []
        for c in range(6):
            k = int(ranks[b, c])
            Ub = U[uo:uo + m[b] * k].reshape(k, m[b]).T   # stored column-major (m x k)
            Vb = V[vo:vo + k * n[b]].reshape(k, n[b])     # row-major (k x n)
            uo += m[b] * k
            vo += k * n[b]
            per.append((Ub, Vb))
        out.append(per)
    st = {"sweeps": int(stats[0]), "launches": int(stats[1]), "pairs_evaluated": float(stats[2]),
          "gpu_eval_s": float(stats[3]), "host_s": float(stats[4]), "max_rank": int(stats[5]),
          "unconverged": int(stats[6]), "pivoting": "device" if device else "host",
          "redone_blocks": int(dinfo[0]),
          "phase_s": {"requests": float(phases[0]), "evaluation": float(phases[1]), "aca_logic": float(phases[2]),
                      "recompression": float(phases[3])}}
    if device:
        st["device_s"] = {"setup": float(dinfo[1]), "sweeps": float(dinfo[2]), "readback": float(dinfo[3])}
    return out, st


def recompress(U, V, eps):
    """Truncate U V to the smallest rank whose discarded singular values have
    Frobenius norm <= eps * ||U V||_F (QR of both factors, SVD of the core,
    the singular values split evenly between the factors), natively on the
    host (``pq3d_lowrank_recompress``; the reference's SVD_recompress,
    Hmatrix.py:116-141, with a relative tolerance)."""
    U = np.ascontiguousarray(np.asarray(U, dtype=np.float64))
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float64))
    m, k = U.shape
    n = V.shape[1]
    if k == 0:
        return U.reshape(m, 0), V.reshape(0, n)
    Uc = np.ascontiguousarray(U.T)                 # column-major m x k
    uo = np.empty(m * k)
    vo = np.empty(k * n)
    r = ctypes.c_longlong()
    check(_lib.lib().pq3d_lowrank_recompress(m, n, k, ptr(Uc), ptr(V), float(eps), ctypes.byref(r), ptr(uo),
                                             ptr(vo)), "pq3d_lowrank_recompress")
    r = r.value
    return uo[:m * r].reshape(r, m).T.copy(), vo[:r * n].reshape(r, n).copy()


# ----------------------------------------------------------------- block tree
class BlockTree:
    """Block tree with the reference's attributes and methods (Hmatrix.py:258-2070)."""

    aca_device = True  # ACA pivot logic and recompression on the GPU (False: pivot logic on host threads)

    def __init__(self, root_block, nodelst, elelst, eleVec, mu_, lambda_, xg, halfspace_jud, mini_leaf,
                 cluster_root=None):
        self.root_block = root_block
        self.nodelst = nodelst
        self.elelst = elelst
        self.eleVec = np.asarray(eleVec, dtype=np.float64)
        self.xg = np.asarray(xg, dtype=np.float64)
        self.mu_ = mu_
        self.lambda_ = lambda_
        self.halfspace_jud = halfspace_jud
        self.mini_leaf = mini_leaf
        self.thread_svd = mini_leaf
        self.blocks_to_process = []
        self.size = 0.0
        self.yvector = np.zeros(len(self.xg))
        self.cluster_root = cluster_root
        self.perm = leaf_order(cluster_root) if cluster_root is not None else None
        self.aca_eps = 1e-4
        self.recompress = True
        # stopping rule of the ACA: "relative" (||u|| ||v|| <= aca_eps ||S_k||_F
        # and small reference residuals, then recompression to aca_eps), or
        # "reference": the reference's absolute rule (Hmatrix.py:797,832-838;
        # stop after a term with pivot < 1e-4 or step ||u|| ||v|| < aca_eps_abs,
        # 0.01 as ACA_worker passes it, Pa per metre of slip) and no
        # recompression, as ACA_worker (:1292-1344) leaves the factors
        self.aca_stop = "relative"
        self.aca_eps_abs = 0.01
        self.build_stats = {}
        self._src = None
        self._device = {}

    # ---- entry evaluation (Hmatrix.py:329-658)
    @property
    def src(self):
        if self._src is None:
            self._src = GFSource(self.nodelst, self.elelst, self.eleVec, self.xg, self.mu_, self.lambda_,
                                 self.halfspace_jud)
        return self._src

    def GetTtstress(self, Stress, col_cluster):
        """Traction (Tt1, Tt2, Tn) on the frames of elements col_cluster from
        stress columns [xx, yy, zz, xy, xz, yz] (Hmatrix.py:306-324)."""
        ev = self.eleVec[np.asarray(col_cluster)]
        e1, e2, nv = ev[:, 0:3], ev[:, 3:6], ev[:, 6:9]
        S = np.asarray(Stress)
        t = np.stack([S[0] * nv[:, 0] + S[3] * nv[:, 1] + S[4] * nv[:, 2],
                      S[3] * nv[:, 0] + S[1] * nv[:, 1] + S[5] * nv[:, 2],
                      S[4] * nv[:, 0] + S[5] * nv[:, 1] + S[2] * nv[:, 2]], axis=1)
        return np.stack([np.sum(t * e1, 1), np.sum(t * e2, 1), np.sum(t * nv, 1)], axis=1)

    def _sub(self, row_cluster, col_cluster):
        return self.src.segments([(row_cluster, col_cluster)])[0]

    def calc_stressfunc_fromCS(self, row_cluster, col_cluster):
        """Strike-slip sub-blocks (A1s, A2s, Bs) over rows x columns (:477-532)."""
        g = self._sub(row_cluster, col_cluster)
        return g[0], g[1], g[2]

    def calc_stressfunc_fromCD(self, row_cluster, col_cluster):
        """Dip-slip sub-blocks (A1d, A2d, Bd) (:536-591)."""
        g = self._sub(row_cluster, col_cluster)
        return g[3], g[4], g[5]

    calc_stressfunc_fromC_rowS = calc_stressfunc_fromCS
    calc_stressfunc_fromC_rowD = calc_stressfunc_fromCD

    def calc_rowsS(self, Istart, Iend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCS(row_cluster[Istart:Iend], col_cluster)

    def calc_rowsD(self, Istart, Iend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCD(row_cluster[Istart:Iend], col_cluster)

    def calc_colsS(self, Jstart, Jend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCS(row_cluster, col_cluster[Jstart:Jend])

    def calc_colsD(self, Jstart, Jend, row_cluster, col_cluster):
        return self.calc_stressfunc_fromCD(row_cluster, col_cluster[Jstart:Jend])

    # ---- ACA (Hmatrix.py:661-1344)
    def _aca_args(self, eps=None):
        """(eps argument, recompression tolerance) for the stopping rule in
        use: a negative eps selects the reference's absolute rule."""
        if self.aca_stop == "reference":
            return -abs(self.aca_eps_abs if eps is None else eps), 0.0
        if self.aca_stop != "relative":
            raise ValueError(f"aca_stop must be 'relative' or 'reference', not {self.aca_stop!r}")
        e = self.aca_eps if eps is None else eps
        return e, (e if self.recompress else 0.0)

    def _aca_one(self, row_cluster, col_cluster, eps, comps):
        e, rc = self._aca_args(eps)
        out, _ = aca_blocks(self.src, [(row_cluster, col_cluster)], eps=e, recompress_eps=rc)
        d = {}
        for c in comps:
            U, V = out[0][c]
            d["U_ACA_" + MATRIX_NAMES[c]] = U
            d["V_ACA_" + MATRIX_NAMES[c]] = V
        return d

    def ACA_plus_full_traceS(self, n_rows, n_cols, calc_rows, calc_cols, row_cluster, col_cluster, eps=None,
                             max_iter=None, verbose=False):
        """Low-rank factors of A1s, A2s, Bs over the block (keys U_ACA_*, V_ACA_*).
        Entries are evaluated on the GPU; ``calc_rows``/``calc_cols`` are
        accepted for signature compatibility.  ``eps`` is the relative
        tolerance (default: the tree's ``aca_eps``), or with
        ``aca_stop = "reference"`` the reference's absolute step tolerance
        (default ``aca_eps_abs``)."""
        return self._aca_one(row_cluster, col_cluster, eps, (0, 1, 2))

    def ACA_plus_full_traceD(self, n_rows, n_cols, calc_rows, calc_cols, row_cluster, col_cluster, eps=None,
                             max_iter=None, verbose=False):
        """Low-rank factors of A1d, A2d, Bd (see ``ACA_plus_full_traceS``)."""
        return self._aca_one(row_cluster, col_cluster, eps, (3, 4, 5))

    def _is_lowrank(self, block):
        return len(block.row_cluster) > self.thread_svd and len(block.col_cluster) > self.thread_svd

    def ACA_worker(self, block):
        """Fill one leaf: low-rank factors if both clusters are larger than
        ``mini_leaf``, else the six dense sub-blocks (Hmatrix.py:1292-1344)."""
        self._fill([block])
        return block

    def collect_blocks(self, block):
        """Append the leaves under ``block`` to ``blocks_to_process`` in
        depth-first order (Hmatrix.py:1618-1624)."""
        stack = [block]
        while stack:
            b = stack.pop()
            if b.is_leaf():
                self.blocks_to_process.append(b)
            else:
                stack.extend(reversed(b.children))

    def calc_Hmatrix_ACA(self, dir=None, save_corefunc=False):
        """Compress every leaf (Hmatrix.py:1599-1614): one batched ACA over all
        low-rank leaves, one batched evaluation of all dense leaves.  With
        ``save_corefunc`` the leaves are written to
        ``dir/blocks_to_process.joblib`` with ``joblib.dump``, as the reference
        does (Hmatrix.py:1552-1553, :1612); this module never loads that file
        (it is a pickle)."""
        if not self.blocks_to_process:
            self.collect_blocks(self.root_block)
        self._fill(self.blocks_to_process)
        if save_corefunc:
            self.save_blocks(dir)
        return self.blocks_to_process

    def save_blocks(self, dir):
        """``dir/blocks_to_process.joblib``: the leaves with their factors or
        dense sub-blocks and flags, as plain arrays and lists."""
        from joblib import dump
        os.makedirs(dir, exist_ok=True)
        path = os.path.join(dir, "blocks_to_process.joblib")
        dump(self.blocks_to_process, path)
        return path

    def _fill(self, blocks):
        lowrank = [b for b in blocks if self._is_lowrank(b)]
        dense = [b for b in blocks if not self._is_lowrank(b)]
        t0 = time.perf_counter()
        if lowrank:
            e, rc = self._aca_args()
            out, st = aca_blocks(self.src, [(b.row_cluster, b.col_cluster) for b in lowrank], eps=e,
                                 recompress_eps=rc, device=self.aca_device)
            t1 = time.perf_counter()
            for b, per in zip(lowrank, out):
                dS, dD = {}, {}
                for c, (U, V) in enumerate(per):
                    d = dS if c < 3 else dD
                    d["U_ACA_" + MATRIX_NAMES[c]] = U
                    d["V_ACA_" + MATRIX_NAMES[c]] = V
                b.ACA_dictS, b.ACA_dictD = dS, dD
                b.judaca = True
                b.judproc = True
                self.size += sum(U.nbytes + V.nbytes for U, V in per) / 2 ** 20
            st["unpack_s"] = time.perf_counter() - t1
            st["aca_s"] = t1 - t0
            self.build_stats["aca"] = st
        t2 = time.perf_counter()
        chunk = 4096
        for s in range(0, len(dense), chunk):
            part = dense[s:s + chunk]
            subs = self.src.segments([(b.row_cluster, b.col_cluster) for b in part])
            for b, g in zip(part, subs):
                # (m, n) views in column-major order, as evaluated and as the
                # device packing reads them: no host copy
                for c, name in enumerate(MATRIX_NAMES):
                    setattr(b, "Mf_" + name, g[c])
                b.judaca = False
                b.judproc = True
                self.size += g.nbytes / 2 ** 20
        self.build_stats["dense_s"] = time.perf_counter() - t2

    # ---- products (Hmatrix.py:1773-1994)
    def device(self, blocks=None, factor_type=_lib.MAT_F32):
        """The device-resident H-matrix of ``blocks`` (default: all leaves)."""
        blocks = self.blocks_to_process if blocks is None else blocks
        key = (id(blocks), len(blocks), factor_type)
        dev = self._device.get(key)
        if dev is None:
            dev = DeviceHmatrix(len(self.xg), self.perm, blocks, factor_type)
            self._device[key] = dev
        return dev

    def blocks_process_MVM(self, xvector, blocks_to_process, type):
        """y = A_type x over the given leaves (Hmatrix.py:1773-1866), computed
        by the device H-matrix product (FP32 factors and dense leaves)."""
        c = MATRIX_NAMES.index(type)
        x = np.asarray(xvector, dtype=np.float64)
        z = np.zeros_like(x)
        y = self.device(blocks_to_process).mvm(x if c < 3 else z, z if c < 3 else x)
        return y[c % 3]

    blocks_process_MVM_gpu = blocks_process_MVM

    def close(self):
        for d in self._device.values():
            d.close()
        self._device = {}
        if self._src is not None:
            self._src.close()
            self._src = None


# ----------------------------------------------------------------- device H-matrix
class DeviceHmatrix:
    """Device-resident H-matrix of the six traction matrices.

    Rows and columns live in the cluster (depth-first leaf) order, so every
    block is a contiguous sub-block.  Dense leaves are stored in FP32, low-rank
    factors in FP32 or BF16 (``factor_type``).  ``mvm(x1, x2)`` returns
    (A1s x1 + A1d x2, A2s x1 + A2d x2, Bs x1 + Bd x2), the three traction rates
    of one step (QDsim.py:1153-1160)."""

    def __init__(self, n, perm, blocks, factor_type=_lib.MAT_F32):
        self.n = int(n)
        perm = _i32(perm)
        nb = len(blocks)
        rows = [_i32(b.row_cluster) for b in blocks]
        cols = [_i32(b.col_cluster) for b in blocks]
        roff = _i64(np.concatenate([[0], np.cumsum([len(r) for r in rows])]))
        coff = _i64(np.concatenate([[0], np.cumsum([len(c) for c in cols])]))
        ridx = np.concatenate(rows)
        cidx = np.concatenate(cols)
        kind = _i32([1 if b.judaca else 0 for b in blocks])
        ranks = np.zeros((nb, 6), dtype=np.int64)
        dparts, uparts, vparts = [], [], []
        for k, b in enumerate(blocks):
            if b.judaca:
                for c, name in enumerate(MATRIX_NAMES):
                    U, V = b.factors(name)
                    U = np.asarray(U, dtype=np.float64).reshape(len(b.row_cluster), -1)
                    V = np.asarray(V, dtype=np.float64).reshape(-1, len(b.col_cluster))
                    ranks[k, c] = U.shape[1]
                    uparts.append(U.T.ravel())   # column-major m x k
                    vparts.append(V.ravel())     # row-major k x n
            else:
                for name in MATRIX_NAMES:
                    M = np.asarray(getattr(b, "Mf_" + name), dtype=np.float64)
                    dparts.append(M.T.ravel())   # column-major m x n
        dense = np.concatenate(dparts) if dparts else np.zeros(1)
        U = np.concatenate(uparts) if uparts else np.zeros(1)
        V = np.concatenate(vparts) if vparts else np.zeros(1)
        h = ctypes.c_void_p()
        check(_lib.lib().pq3d_hmat_create(ctypes.byref(h), self.n, _ip(perm), nb, _lp(roff), _ip(ridx), _lp(coff),
                                          _ip(cidx), _ip(kind), _lp(_i64(ranks)), ptr(dense), ptr(U), ptr(V),
                                          int(factor_type)), "pq3d_hmat_create")
        self.h = h
        self.factor_type = factor_type

    def info(self):
        o = np.zeros(10)
        check(_lib.lib().pq3d_hmat_info(self.h, ptr(o)), "pq3d_hmat_info")
        return {"dense_bytes": o[0], "factor_bytes": o[1], "bytes_per_mvm": o[2], "bytes_per_mvm_fixtn": o[3],
                "segments": int(o[4]), "row_waves": int(o[5]), "col_waves": int(o[6]), "lowrank_blocks": int(o[7]),
                "stored_bytes_per_mvm": o[8], "stored_bytes_per_mvm_fixtn": o[9]}

    def mvm(self, x1, x2, fix_tn=False):
        x1 = np.ascontiguousarray(x1, dtype=np.float64)
        x2 = np.ascontiguousarray(x2, dtype=np.float64)
        y = np.empty(3 * self.n)
        check(_lib.lib().pq3d_hmat_mvm(self.h, ptr(x1), ptr(x2), int(bool(fix_tn)), ptr(y)), "pq3d_hmat_mvm")
        return y.reshape(3, self.n)

    def time_mvm(self, reps=20, fix_tn=False):
        f = ctypes.c_float()
        check(_lib.lib().pq3d_hmat_time_mvm(self.h, int(reps), int(bool(fix_tn)), ctypes.byref(f)),
              "pq3d_hmat_time_mvm")
        return f.value

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().pq3d_hmat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build(nodes, tri, mu, lam, halfspace=True, mini_leaf=32, min_size=16, eta=2.0, eps=1e-4,
          recompress_factors=True, aca_device=True):
    """Mesh -> compressed ``BlockTree`` with ``blocks_to_process`` filled.
    ``nodes`` in metres, ``tri`` 1-based (the reader's conventions);
    aca_device: pivot logic and recompression of the ACA on the GPU (False:
    pivot logic on host threads)."""
    eleVec, xg = readmsh.get_eleVec(nodes, tri, False)
    para = {"Shear modulus": mu, "Lame constants": lam, "Half space": bool(halfspace)}
    t0 = time.perf_counter()
    tree = createHmatrix(xg, nodes, tri, eleVec, para, mini_leaf=mini_leaf, min_size=min_size, eta=eta)
    tree.collect_blocks(tree.root_block)
    tree.build_stats["tree_s"] = time.perf_counter() - t0
    tree.aca_eps = eps
    tree.recompress = recompress_factors
    tree.aca_device = aca_device
    tree.calc_Hmatrix_ACA()
    return tree
