"""Edge cases of the device paths (GPU).

* the ctypes drop-ins with zero receivers / zero sources return without
  touching the outputs and leave no error;
* a one-element mesh (self-interaction only): the dense step sequence
  against the NumPy FP64 oracle (bars of the dense tests: 1e-9), and the
  ensemble GEMM against the FP64 product of the FP32 matrix (2e-6 of the
  magnitude sum) plus every member's steps against the oracle (1e-6);
* a mesh size that is not a multiple of any tile (n = 70): dense matvec per
  row within 1e-12 of sum |a_ij u_j| (FP64 storage);
* ACA with no blocks, and with blocks smaller than a wave, on the device
  pivoting path.
"""
import ctypes

import numpy as np
import pytest

from oracle import qdsim_oracle as Q

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9


def test_dropins_with_no_points_are_noops():
    from pyquake3d_amd import _lib
    L = _lib.lib()
    d = ctypes.POINTER(ctypes.c_double)
    st = (ctypes.c_double * 6)(*([7.0] * 6))
    sn = (ctypes.c_double * 6)(*([7.0] * 6))
    p = (ctypes.c_double * 3)(0.0, 0.0, -1000.0)
    q = (ctypes.c_double * 3)(1000.0, 0.0, -1000.0)
    r = (ctypes.c_double * 3)(0.0, 0.0, -2000.0)
    L.TDstressFS_C(None, None, None, 0, p, q, r, 1.0, 0.0, 0.0, MU, LAM, True, st, sn)
    L.TDstressEachSourceAtReceiver_C(0.0, 500.0, -1500.0, None, None, None, 0, 1.0, 0.0, 0.0, MU, LAM, True, st,
                                     sn)
    assert list(st) == [7.0] * 6 and list(sn) == [7.0] * 6
    assert L.pq3d_last_error() in (b"", None)


def _one_element():
    from pyquake3d_amd import greenfunc, readmsh, synth
    nodes_km = np.array([[0.0, 0.0, -5.0], [2.0, 0.0, -5.0], [0.0, 0.0, -7.0]])
    tri = np.array([[1, 2, 3]])
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    cols = synth.friction_params(xg, seed=4, rake=0.5)
    cols[:, 1] = 0.004   # velocity weakening, so the element evolves
    return nodes, tri, ev, xg, cols, greenfunc.source_vertices(nodes, tri)


def test_single_element_dense_steps_follow_oracle():
    from pyquake3d_amd import _lib
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes, tri, ev, xg, cols, P = _one_element()
    s = DenseSim(1, _lib.MAT_F64)
    s.assemble(*P, xg, ev, MU, LAM, True)
    orc = Q.from_init_columns([s.download_matrix(k) for k in range(6)], cols, V0, VPL, MU, CS, True)
    for name in ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "a", "b", "dc", "f0",
                 "shear_loading", "normal_loading"):
        s.set(name, getattr(orc, name))
    s.set("rake0", orc.rake0)
    s.set_params(V0, VPL, MU / (2 * CS), False, True)
    dt = 1e-3
    for k in range(60):
        h, dtn, nr = s.step(dt)
        ho, dto = orc.step(dt)
        assert abs(h - ho) <= 1e-9 * ho and abs(dtn - dto) <= 1e-9 * dto, (k, h, ho)
        dt = dtn
    for name in ("state", "Tt1o", "slip"):
        got, want = s.get(name), getattr(orc, name)
        assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max(), name
    s.close()


def test_single_element_ensemble():
    from pyquake3d_amd import _lib, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    from pyquake3d_amd.ensemble import Ensemble, init_members
    nodes, tri, ev, xg, cols, P = _one_element()
    s = DenseSim(1, _lib.MAT_F32)
    s.assemble(*P, xg, ev, MU, LAM, True)
    mats = [s.download_matrix(k) for k in range(6)]
    mcols = np.stack([synth.member_params(xg, m) for m in range(64)])
    e = Ensemble(1, 64, matrices_from=s)
    init_members(e, mcols, V0, VPL, MU, CS)
    e.gemm()
    got = e.gemm_result()
    u1 = e.get("slipv1") - VPL * np.cos(e.get("rake0"))
    u2 = e.get("slipv2") - VPL * np.sin(e.get("rake0"))
    u1, u2 = u1.astype(np.float32).astype(np.float64), u2.astype(np.float32).astype(np.float64)
    for q in range(3):
        want = u1 @ mats[q].T + u2 @ mats[q + 3].T
        mag = np.abs(u1) @ np.abs(mats[q]).T + np.abs(u2) @ np.abs(mats[q + 3]).T
        assert (np.abs(got[q] - want) / np.where(mag > 0, mag, 1.0)).max() < 2e-6, q
    e.run(12, graph_steps=4)
    mem = e.members()
    assert np.all(mem["status"] == 0)
    states = e.get("state")
    for m in (0, 31, 63):
        orc = Q.from_init_columns(mats, mcols[m], V0, VPL, MU, CS)
        dt = 1e-3
        for _ in range(int(mem["steps"][m])):
            h, dt = orc.step(dt)
        assert abs(mem["time"][m] - orc.time) <= 1e-6 * orc.time
        assert np.abs(states[m] - orc.state).max() <= 1e-6 * np.abs(orc.state).max()
    e.close()
    s.close()


def test_ragged_size_dense_product():
    from pyquake3d_amd import _lib, greenfunc, readmsh, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.planar_fault_mesh(7, 5, 14.0, 10.0)
    tri = tri[:70]
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    s = DenseSim(len(tri), _lib.MAT_F64, kc=32)
    s.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    rng = np.random.default_rng(2)
    v1, v2, r0 = rng.normal(size=70) * 1e-6, rng.normal(size=70) * 1e-6, rng.uniform(0, 1, 70)
    for k, v in (("slipv1", v1), ("slipv2", v2), ("rake0", r0), ("normal_loading", np.zeros(70))):
        s.set(k, v)
    s.set_params(V0, VPL, MU / (2 * CS), False, True)
    s.matvec()
    mats = [s.download_matrix(k) for k in range(6)]
    u1, u2 = v1 - VPL * np.cos(r0), v2 - VPL * np.sin(r0)
    for q, name in enumerate(("AdotV1", "AdotV2", "dsigmadt")):
        want = mats[q] @ u1 + mats[q + 3] @ u2
        mag = np.abs(mats[q]) @ np.abs(u1) + np.abs(mats[q + 3]) @ np.abs(u2)
        got = s.get(name)
        assert (np.abs(got - want) / np.where(mag > 0, mag, 1.0)).max() < 1e-12, name
    s.close()


def test_aca_with_no_blocks_and_blocks_smaller_than_a_wave():
    """No blocks: an empty result on both paths.  Blocks of 3 x 5 and 2 x 2
    elements (fewer rows than a wave's lanes) and a well-separated 40 x 40:
    the device pivoting meets the ACA bar of the H-matrix tests, with ranks
    equal to the host loop's."""
    from pyquake3d_amd import Hmatrix as H, readmsh, synth
    nodes_km, tri = synth.planar_fault_mesh(20, 10, 40.0, 20.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    src = H.GFSource(nodes, tri, ev, xg, MU, LAM, True)
    for device in (False, True):
        out, st = H.aca_blocks(src, [], eps=1e-4, device=device)
        assert out == [] and st["unconverged"] == 0
    n = len(tri)
    blocks = [(np.arange(3), np.arange(n - 5, n)), (np.array([7, 9]), np.array([n - 1, n - 3])),
              (np.arange(40), np.arange(n - 40, n))]
    dense = src.segments(blocks)
    host, _ = H.aca_blocks(src, blocks, eps=1e-4)
    dev, st = H.aca_blocks(src, blocks, eps=1e-4, device=True)
    assert st["pivoting"] == "device"
    for A, ph, pd in zip(dense, host, dev):
        strongest = max(np.linalg.norm(A[c]) for c in range(6))
        for c in range(6):
            U, V = pd[c]
            err = np.linalg.norm(A[c] - U @ V)
            assert err <= 10 * 1e-4 * np.linalg.norm(A[c]) + 1e-8 * strongest, (A[c].shape, c, err)
            assert U.shape[1] == ph[c][0].shape[1], (A[c].shape, c)
    src.close()
