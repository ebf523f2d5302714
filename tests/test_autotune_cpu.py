"""Host logic of the split-K autotuner (no GPU): candidate sets, the sweep's
choice, and the JSON cache round trip."""
import json

import pytest

from pyquake3d_amd import autotune


class FakeTiled:
    """Stands in for an Ensemble / DenseSim: the width is rounded like the
    ensemble's (multiple of 4) and the time is a convex function of it."""

    def __init__(self, n, best):
        self.n, self.B, self._best, self._kc = n, 64, best, 256
        self.calls = 0
        self.mat_type, self.fix_tn, self.has_dense = 1, False, True

    @property
    def kc(self):
        return self._kc

    def set_kc(self, kc):
        self._kc = (kc + 3) // 4 * 4

    def time_product(self, reps):
        self.calls += 1
        return 1.0 + ((self._kc - self._best) / self.n) ** 2


@pytest.fixture
def cache(tmp_path, monkeypatch):
    p = tmp_path / "tune.json"
    monkeypatch.setenv("PQ3D_TUNE_CACHE", str(p))
    monkeypatch.setattr(autotune, "_device_key", lambda: "gfx950-test|cu256")
    return p


def test_ensemble_candidates_cover_chunk_counts():
    for n in (37, 2000, 20000, 30000):
        c = autotune.ensemble_candidates(n)
        assert all(k % 4 == 0 and k >= 4 for k in c)
        assert len(set(c)) == len(c)
        chunks = sorted({-(-n // k) for k in c})
        assert chunks[0] == 1 and max(chunks) <= 32
        if n >= 32 * 4:
            assert len(chunks) >= 24  # nearly every chunk count of 1..32 is reachable


def test_dense_candidates():
    assert autotune.dense_candidates(100) == [32, 64, 96, 100]
    c = autotune.dense_candidates(20000)
    assert c[-1] == 20000 and c == sorted(c)


def test_sweep_picks_and_applies_the_fastest(cache):
    obj = FakeTiled(20000, best=2224)
    res = autotune.tune_ensemble(obj)
    assert res["kc"] == min(res["table"], key=lambda t: t[1])[0]   # deterministic times: finalists agree
    assert obj.kc == res["kc"]
    assert abs(res["kc"] - 2224) <= 2224 // 8   # the candidate nearest the optimum
    d = json.loads(cache.read_text())
    assert d[res["key"]]["kc"] == res["kc"]


def test_cached_width_is_applied_without_measuring(cache):
    obj = FakeTiled(20000, best=5000)
    kc = autotune.ensure_ensemble(obj)
    first = obj.calls
    assert first > 0 and kc == obj.kc
    other = FakeTiled(20000, best=5000)
    assert autotune.ensure_ensemble(other) == kc
    assert other.calls == 0
    dense = FakeTiled(20000, best=512)
    assert autotune.ensure_dense(dense) == 512
    assert autotune.cached(autotune.dense_key(dense)) == 512


def test_unwritable_cache_still_tunes(tmp_path, monkeypatch):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    monkeypatch.setenv("PQ3D_TUNE_CACHE", str(blocker / "sub" / "tune.json"))
    monkeypatch.setattr(autotune, "_device_key", lambda: "gfx950-test|cu256")
    obj = FakeTiled(4000, best=1000)
    assert autotune.tune_ensemble(obj)["kc"] == obj.kc


def test_switch_off(monkeypatch):
    monkeypatch.setenv("PQ3D_AUTOTUNE", "0")
    assert not autotune.enabled()
    monkeypatch.setenv("PQ3D_AUTOTUNE", "1")
    assert autotune.enabled()


class FakeSplit(FakeTiled):
    """An ensemble with the BF16 split requested: widths round to 32; the
    split runs only when `active` (Bs = Bd = 0)."""

    def __init__(self, n, best, active):
        super().__init__(n, best)
        self._active = active

    def set_kc(self, kc):
        self._kc = (kc + 31) // 32 * 32

    @property
    def precision(self):
        return "bf16x3", self._active

    @property
    def structure(self):
        return (0b100100, True) if self._active else (0, False)


def test_split_widths_and_cache_keys(cache):
    on, off = FakeSplit(20000, 3360, True), FakeSplit(20000, 3360, False)
    c = autotune.ensemble_candidates(20000, 32)
    assert all(k % 32 == 0 for k in c) and len(set(c)) == len(c)
    r_on, r_off = autotune.tune_ensemble(on), autotune.tune_ensemble(off)
    # both sweep 32-multiples (no duplicate candidates after rounding) ...
    assert all(k % 32 == 0 for k, _ in r_on["table"]) and len({k for k, _ in r_on["table"]}) == len(r_on["table"])
    assert all(k % 32 == 0 for k, _ in r_off["table"]) and len({k for k, _ in r_off["table"]}) == len(r_off["table"])
    # ... but only the running split GEMM gets its own cache entry
    assert r_on["key"].endswith("|nq2|bf16x3") and r_off["key"].endswith("|nq3")
    d = json.loads(cache.read_text())
    assert d[r_on["key"]]["kc"] == r_on["kc"] and d[r_off["key"]]["kc"] == r_off["kc"]
