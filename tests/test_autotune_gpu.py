"""GPU tests of the split-K re-tiling and the autotuner.

Bars:
* the ensemble GEMM at any split-K width kc, against the FP64 product of the
  same FP32 matrices, relative to sum_j |a_ij u_j|: every entry within
  (kc + 8) u (u = 2^-24; each of the chunk's FP32 accumulations rounds once)
  and the root-mean-square within sqrt(kc) u (rounding errors of random
  sign); the tuned width agrees with the single-chunk product to the sum of
  both entry bounds;
* the dense product (FP64 matrices, FP64 sums) at widths 32 and n: traction
  rates agree to 1e-13 of sum_j |a_ij u_j|;
* stepping after a re-tile: 40 steps at width 64 and at the tuned width give
  step sizes and states within 1e-10 relative;
* the tuner applies the fastest width it measured and caches it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU = LAM = 32038120320.0
CS = 3464.0
V0, VPL = 1e-6, 1e-9
B = 64


@pytest.fixture(scope="module")
def mesh():
    from pyquake3d_amd import readmsh, synth
    nodes_km, tri = synth.planar_fault_mesh(24, 12, 48.0, 24.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    return nodes, tri, ev, xg


def _sim(mesh, mat_type):
    from pyquake3d_amd import greenfunc, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    import bench
    nodes, tri, ev, xg = mesh
    s = DenseSim(len(tri), mat_type)
    s.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    bench.init_state(s, synth.friction_params(xg, dc_sigma=0.2, seed=5, rake=0.3))
    return s


def test_gemm_any_width_matches_fp64(mesh, tmp_path, monkeypatch):
    from pyquake3d_amd import _lib, autotune, synth
    from pyquake3d_amd.ensemble import Ensemble, init_members
    monkeypatch.setenv("PQ3D_TUNE_CACHE", str(tmp_path / "t.json"))
    sim = _sim(mesh, _lib.MAT_F32)
    n = sim.n
    mats = [sim.download_matrix(k) for k in range(6)]
    xg = mesh[3]
    e = Ensemble(n, B, matrices_from=sim, precision="fp32")
    init_members(e, np.stack([synth.member_params(xg, m) for m in range(B)]), V0, VPL, MU, CS)
    rng = np.random.default_rng(9)
    V1 = rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n))
    V2 = rng.lognormal(-18, 3, (B, n))
    e.set("slipv1", V1)
    e.set("slipv2", V2)
    r0 = e.get("rake0")
    U1 = (V1 - VPL * np.cos(r0)).astype(np.float32).astype(np.float64)
    U2 = (V2 - VPL * np.sin(r0)).astype(np.float32).astype(np.float64)
    u = 2.0 ** -24
    want = np.stack([U1 @ mats[q].T + U2 @ mats[q + 3].T for q in range(3)])
    mag = np.stack([np.abs(U1) @ np.abs(mats[q]).T + np.abs(U2) @ np.abs(mats[q + 3]).T for q in range(3)])
    mag = np.where(mag > 0, mag, 1.0)
    results = {}
    for kc in (4, 100, n // 3, n):
        e.set_kc(kc)
        assert e.kc % 4 == 0 and e.kc >= min(kc, n) and e.kc < min(kc, n) + 4
        e.gemm()
        got = e.gemm_result()
        err = np.abs(got - want) / mag
        assert err.max() < (e.kc + 8) * u, (kc, err.max())
        assert np.sqrt(np.mean(err ** 2)) < np.sqrt(e.kc) * u, (kc, np.sqrt(np.mean(err ** 2)))
        results[e.kc] = got
    res = autotune.tune_ensemble(e)
    assert e.kc == res["kc"]
    assert res["kc"] in [k for k, _ in sorted(res["table"], key=lambda t: t[1])[:4]]  # a re-timed finalist
    assert autotune.cached(autotune.ensemble_key(e)) == res["kc"]
    e.gemm()
    got = e.gemm_result()
    full = (n + 3) // 4 * 4
    assert (np.abs(got - results[full]) / mag).max() < (e.kc + full + 16) * u
    e.close()
    sim.close()


def test_split_gemm_any_width_matches_fp64(mesh, tmp_path, monkeypatch):
    """The BF16-split GEMM (three exact BF16 parts per FP32 operand, six
    products, 32-source steps) at widths 32, 128, n / 3 and n, against the
    FP64 product of the same FP32 operands: the FP32 GEMM's bars, (kc + 8) u
    per entry and sqrt(kc) u root-mean-square of sum_j |a_ij u_j|; the
    tuner's widths are multiples of 32 and cached under their own key."""
    from pyquake3d_amd import _lib, autotune, synth
    from pyquake3d_amd.ensemble import Ensemble, init_members
    monkeypatch.setenv("PQ3D_TUNE_CACHE", str(tmp_path / "t.json"))
    sim = _sim(mesh, _lib.MAT_F32)
    n = sim.n
    mats = [sim.download_matrix(k) for k in range(6)]
    xg = mesh[3]
    e = Ensemble(n, B, matrices_from=sim)
    assert e.precision == ("bf16x3", True)
    init_members(e, np.stack([synth.member_params(xg, m) for m in range(B)]), V0, VPL, MU, CS)
    rng = np.random.default_rng(19)
    V1 = rng.lognormal(-18, 3, (B, n)) * rng.choice([-1, 1], (B, n))
    V2 = rng.lognormal(-18, 3, (B, n))
    e.set("slipv1", V1)
    e.set("slipv2", V2)
    r0 = e.get("rake0")
    U1 = (V1 - VPL * np.cos(r0)).astype(np.float32).astype(np.float64)
    U2 = (V2 - VPL * np.sin(r0)).astype(np.float32).astype(np.float64)
    u = 2.0 ** -24
    want = np.stack([U1 @ mats[q].T + U2 @ mats[q + 3].T for q in range(2)])
    mag = np.stack([np.abs(U1) @ np.abs(mats[q]).T + np.abs(U2) @ np.abs(mats[q + 3]).T for q in range(2)])
    mag = np.where(mag > 0, mag, 1.0)
    for kc in (32, 100, n // 3, n):
        e.set_kc(kc)
        assert e.kc % 32 == 0 and e.kc >= min(kc, n) and e.kc < min(kc, n) + 32, (kc, e.kc)
        e.gemm()
        got = e.gemm_result()
        assert not np.any(got[2])   # Bs = Bd = 0: the normal component is exactly zero
        err = np.abs(got[:2] - want) / mag
        assert err.max() < (e.kc + 8) * u, (kc, err.max())
        assert np.sqrt(np.mean(err ** 2)) < np.sqrt(e.kc) * u, (kc, np.sqrt(np.mean(err ** 2)))
    res = autotune.tune_ensemble(e)
    assert res["kc"] % 32 == 0 and all(k % 32 == 0 for k, _ in res["table"])
    assert res["key"].endswith("|bf16x3") and autotune.cached(res["key"]) == res["kc"]
    e.set_precision("fp32")
    assert e.precision == ("fp32", False) and autotune.ensemble_key(e) != res["key"]
    e.close()
    sim.close()


def test_dense_product_any_width(mesh):
    from pyquake3d_amd import _lib
    sim = _sim(mesh, _lib.MAT_F64)
    n = sim.n
    out = {}
    for kc in (32, 256, n):
        sim.set_kc(kc)
        assert sim.kc == kc
        sim.matvec()
        out[kc] = np.stack([sim.get("AdotV1"), sim.get("AdotV2"), sim.get("dsigmadt")])
    mats = [sim.download_matrix(k) for k in range(6)]
    u1 = sim.get("slipv1") - VPL * np.cos(sim.get("rake0"))
    u2 = sim.get("slipv2") - VPL * np.sin(sim.get("rake0"))
    mag = np.stack([np.abs(mats[q]) @ np.abs(u1) + np.abs(mats[q + 3]) @ np.abs(u2) for q in range(3)])
    mag = np.where(mag > 0, mag, 1.0)   # the normal component of a planar fault is zero
    for kc in (32, 256):
        assert (np.abs(out[kc] - out[n]) / mag).max() < 1e-13, kc
    sim.close()


def test_stepping_after_retile(mesh, tmp_path, monkeypatch):
    from pyquake3d_amd import _lib, autotune
    monkeypatch.setenv("PQ3D_TUNE_CACHE", str(tmp_path / "t.json"))
    a = _sim(mesh, _lib.MAT_F32)
    b = _sim(mesh, _lib.MAT_F32)
    a.set_kc(64)
    res = autotune.tune_dense(b)
    assert b.kc == res["kc"]
    ra, rb = [], []
    for s, rec in ((a, ra), (b, rb)):
        dt = 1e-3
        for _ in range(40):
            h, dt, _ = s.step(dt)
            rec.append((h, dt))
    ra, rb = np.array(ra), np.array(rb)
    assert np.allclose(ra, rb, rtol=1e-10, atol=0), np.abs(ra / rb - 1).max()
    sa, sb = a.get("state"), b.get("state")
    assert np.abs(sa - sb).max() <= 1e-10 * np.abs(sa).max()
    a.close()
    b.close()


def test_zero_normal_matrices_skipped_bit_identically(mesh):
    """Planar vertical fault: Bs = Bd = 0 exactly (mirror symmetry).  The scan
    finds them, the dense product and the ensemble GEMM skip them, and every
    result is bit-identical to streaming all six matrices."""
    from pyquake3d_amd import _lib, synth
    from pyquake3d_amd.ensemble import Ensemble, init_members
    a = _sim(mesh, _lib.MAT_F32)
    b = _sim(mesh, _lib.MAT_F32)
    assert not np.any(a.download_matrix(2)) and not np.any(a.download_matrix(5))
    assert a.structure == (0b100100, True)
    b.exploit_structure = False
    assert b.structure == (0, False)
    ra, rb = [], []
    for s, rec in ((a, ra), (b, rb)):
        dt = 1e-3
        for _ in range(20):
            h, dt, _ = s.step(dt)
            rec.append((h, dt))
        st, r = s.run(16, graph_steps=8, records=True)
        assert st == 0
        rec.extend(map(tuple, r[:, :2]))
    assert np.array_equal(np.array(ra), np.array(rb))
    for name in ("state", "Tt1o", "Tt2o", "Tno", "slipv", "dsigmadt"):
        assert np.array_equal(a.get(name), b.get(name)), name
    xg = mesh[3]
    cols = np.stack([synth.member_params(xg, m) for m in range(B)])
    # FP32 products on both sides (the BF16-split GEMM serves only the
    # two-component case, so it has no six-matrix twin to compare with)
    e1 = Ensemble(a.n, B, matrices_from=a, precision="fp32")
    e2 = Ensemble(a.n, B, matrices_from=a, precision="fp32")
    e2.exploit_structure = False
    for e in (e1, e2):
        init_members(e, cols, V0, VPL, MU, CS)
    assert e1.structure[1] and not e2.structure[1]
    e1.gemm()
    e2.gemm()
    assert np.array_equal(e1.gemm_result(), e2.gemm_result())
    e1.run(10, graph_steps=5)
    e2.run(10, graph_steps=5)
    for name in ("state", "Tt1o", "slip", "dsigmadt"):
        assert np.array_equal(e1.get(name), e2.get(name)), name
    for o in (e1, e2, a, b):
        o.close()


def test_rough_fault_streams_all_matrices():
    from pyquake3d_amd import _lib, greenfunc, readmsh, synth
    from pyquake3d_amd.QDsim_gpu import DenseSim
    nodes_km, tri = synth.rough_fault_mesh(12, 6, length_km=24.0, depth_km=12.0)
    nodes = nodes_km * 1e3
    ev, xg = readmsh.get_eleVec(nodes, tri, False)
    s = DenseSim(len(tri), _lib.MAT_F32)
    s.assemble(*greenfunc.source_vertices(nodes, tri), xg, ev, MU, LAM, True)
    assert s.structure == (0, False)
    s.upload_matrix(2, np.zeros((s.n, s.n)))   # one zero matrix is not enough
    assert s.structure == (0b100, False)
    s.close()
