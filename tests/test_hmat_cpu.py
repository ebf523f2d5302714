"""CPU checks of the H-matrix tree layer (no GPU needed).

Pins:
* hand-worked cluster splits: midpoint of the largest extent, ties at the
  midpoint go left (``<= gamma``), first axis on equal extents, no split of
  coincident points;
* admissibility, diameter and box distance on boxes with closed-form values;
* the product's leaf list equals the oracle restatement's (same order, same
  row/column sets, same low-rank/dense decision) on a rough fault mesh;
* partition: every matrix entry lies in exactly one leaf; every low-rank
  leaf is admissible; every cluster is a contiguous range of the cluster
  order;
* recompression (native host code): a rank-5 matrix written with redundant
  terms comes back at rank 5; on a matrix with known singular values the
  kept rank is the one the tolerance asks for and the result equals the
  truncated SVD to 1e-12; the zero matrix has rank 0.
"""
import os

import numpy as np
import pytest

from oracle import hmat_oracle as HO
from pyquake3d_amd import Hmatrix as H

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyquake3d_amd", "libpq3d.so")


def _leaf_sets(node, out):
    if node.left is None:
        out.append(sorted(node.indices.tolist()))
    else:
        _leaf_sets(node.left, out)
        _leaf_sets(node.right, out)
    return out


def _oracle_leaf_sets(nodes, k=0, out=None):
    out = [] if out is None else out
    if nodes[k]["kids"] is None:
        out.append(sorted(nodes[k]["idx"].tolist()))
    else:
        _oracle_leaf_sets(nodes, nodes[k]["kids"][0], out)
        _oracle_leaf_sets(nodes, nodes[k]["kids"][1], out)
    return out


def test_cluster_split_rules_hand_worked():
    line4 = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]])
    # [0, 3] -> gamma 1.5 -> {0, 1} {2, 3} -> singletons
    root = H.build_block_tree(np.arange(4), line4, min_size=1)
    assert root.split_dim == 0 and root.split_value == 1.5
    assert _leaf_sets(root, []) == [[0], [1], [2], [3]]
    # tie at the midpoint goes left: [0, 2] -> gamma 1 -> {0, 1} {2}
    line3 = line4[:3]
    root = H.build_block_tree(np.arange(3), line3, min_size=1)
    assert sorted(root.left.indices.tolist()) == [0, 1] and root.right.indices.tolist() == [2]
    # equal extents in x and y: the first axis (x) is split
    sq = np.array([[0.0, 0, 0], [2, 0, 0], [0, 2, 0], [2, 2, 0]])
    root = H.build_block_tree(np.arange(4), sq, min_size=1)
    assert root.split_dim == 0 and sorted(root.left.indices.tolist()) == [0, 2]
    # coincident points cannot be split: a single leaf above min_size
    same = np.zeros((5, 3))
    root = H.build_block_tree(np.arange(5), same, min_size=1)
    assert root.left is None and len(root.indices) == 5
    # the oracle restatement agrees on all four
    for pts in (line4, line3, sq, same):
        assert _leaf_sets(H.build_block_tree(np.arange(len(pts)), pts, min_size=1), []) == \
            _oracle_leaf_sets(HO.cluster_leaves(pts, min_size=1))


def test_admissibility_closed_forms():
    a = np.array([[0.0, 0, 0], [1, 1, 1]])                  # unit cube
    b = np.array([[3.0, 0, 0], [4, 1, 1]])                  # shifted by 3 in x: gap 2
    pts = np.vstack([a, b])
    assert H.diameter(a[0], a[1]) == pytest.approx(np.sqrt(3.0), rel=0, abs=1e-15)
    assert H.distance(a[0], a[1], b[0], b[1]) == pytest.approx(2.0, rel=0, abs=1e-15)
    assert H.distance(b[0], b[1], a[0], a[1]) == pytest.approx(2.0, rel=0, abs=1e-15)
    assert H.is_admissible([0, 1], [2, 3], pts, eta=2.0)        # sqrt(3) <= 4
    assert not H.is_admissible([0, 1], [2, 3], pts, eta=0.5)    # sqrt(3) > 1
    # diagonal gap: boxes apart by 3 in x and 4 in y -> distance 5
    c = np.array([[4.0, 5, 0], [5, 6, 1]])
    assert H.distance(a[0], a[1], c[0], c[1]) == pytest.approx(5.0, rel=0, abs=1e-15)
    # overlapping boxes are never admissible unless one is a point
    d = np.array([[0.5, 0.5, 0.5], [2, 2, 2]])
    assert H.distance(a[0], a[1], d[0], d[1]) == 0.0
    assert not H.is_admissible([0, 1], [2, 3], np.vstack([a, d]))
    assert HO.admissible(a[0], a[1], b[0], b[1], 2.0) and not HO.admissible(a[0], a[1], b[0], b[1], 0.5)
    assert HO.admissible(a[0], a[1], c[0], c[1], 2.0)


@pytest.fixture(scope="module")
def rough_tree():
    from pyquake3d_amd import readmsh, synth
    nodes_km, tri = synth.rough_fault_mesh(40, 25, length_km=16.0, depth_km=7.5, seed=3)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    para = {"Shear modulus": 3.2e10, "Lame constants": 3.2e10, "Half space": True}
    tree = H.createHmatrix(xg, nodes, tri, ev, para, mini_leaf=32)
    tree.collect_blocks(tree.root_block)
    return tree, xg


def test_leaves_match_oracle_restatement(rough_tree):
    tree, xg = rough_tree
    got = [(sorted(b.row_cluster.tolist()), sorted(b.col_cluster.tolist()), tree._is_lowrank(b))
           for b in tree.blocks_to_process]
    want = [(sorted(r.tolist()), sorted(c.tolist()), lr) for r, c, lr in HO.block_leaves(xg, 16, 32, 2.0)]
    assert len(got) == len(want)
    assert got == want
    assert any(lr for _, _, lr in got) and not all(lr for _, _, lr in got)


def test_partition_admissibility_and_cluster_order(rough_tree):
    tree, xg = rough_tree
    n = len(xg)
    cover = np.zeros((n, n), dtype=np.int32)
    pos = np.empty(n, dtype=np.int64)
    pos[tree.perm] = np.arange(n)
    assert np.array_equal(np.sort(tree.perm), np.arange(n))
    for b in tree.blocks_to_process:
        cover[np.ix_(b.row_cluster, b.col_cluster)] += 1
        for ids, node in ((b.row_cluster, b.row_node), (b.col_cluster, b.col_node)):
            p = pos[ids]
            assert p.min() == node.start and p.max() - p.min() + 1 == len(ids)
        if tree._is_lowrank(b):
            assert H.is_admissible(b.row_cluster, b.col_cluster, xg, eta=2.0)
    assert cover.min() == 1 and cover.max() == 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpq3d.so not built")
def test_recompress_recovers_rank():
    rng = np.random.default_rng(1)
    X, Y = rng.standard_normal((60, 5)), rng.standard_normal((5, 50))
    A = X @ Y
    # redundant factors: the rank-5 part split over 10 terms, plus a 1e-12 term
    U = np.hstack([X, X, 1e-12 * rng.standard_normal((60, 1))])
    V = np.vstack([0.25 * Y, 0.75 * Y, rng.standard_normal((1, 50))])
    u, v = H.recompress(U, V, 1e-8)
    assert u.shape[1] == 5 and v.shape[0] == 5
    assert np.linalg.norm(u @ v - A) <= 1e-8 * np.linalg.norm(A)
    u0, v0 = H.recompress(np.zeros((7, 0)), np.zeros((0, 9)), 1e-4)
    assert u0.shape == (7, 0) and v0.shape == (0, 9)
    # against the SVD: a matrix with singular values 10^-j keeps the ranks the
    # tolerance asks for, and the kept part is the truncated SVD
    qa, _ = np.linalg.qr(rng.standard_normal((40, 8)))
    qb, _ = np.linalg.qr(rng.standard_normal((30, 8)))
    s = 10.0 ** -np.arange(8)
    A = (qa * s) @ qb.T
    for eps, want in ((1e-2, 2), (1e-5, 5)):
        u, v = H.recompress(qa * s, qb.T, eps)
        assert u.shape[1] == want
        best = (qa[:, :want] * s[:want]) @ qb[:, :want].T
        assert np.linalg.norm(u @ v - best) <= 1e-12 * np.linalg.norm(A)
    # the exact zero matrix has rank 0
    u, v = H.recompress(np.zeros((5, 3)), np.ones((3, 4)), 1e-4)
    assert u.shape == (5, 0) and v.shape == (0, 4)
