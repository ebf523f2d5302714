"""The three-part BF16 split of the ensemble GEMM (ens_kernels.hip, split3f),
restated in NumPy: hi = x truncated to its top 8 significant bits (the high
half of the FP32 pattern), r = x - hi, mid = r truncated likewise, lo = r - mid.
Checked on random FP32 values over 60 binades and on edge values: the
residuals are exact in FP32, x = hi + mid + lo exactly, every part has at
most 8 significant bits (so its high half-word is the part itself, as the
kernel packs it), and |mid| <= 2^-7 |x|, |lo| <= 2^-15 |x|: the six kept
products leave out terms of at most 3 * 2^-22 |a u| in total."""
import numpy as np

MASK = np.uint32(0xFFFF0000)


def _trunc(x):
    return (x.view(np.uint32) & MASK).view(np.float32)


def _split(x):
    x = np.asarray(x, dtype=np.float32)
    hi = _trunc(x)
    r = (x - hi).astype(np.float32)
    mid = _trunc(r)
    lo = (r - mid).astype(np.float32)
    return hi, r, mid, lo


def test_split_is_exact_and_parts_are_bf16():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(200000) * np.exp2(rng.integers(-30, 30, 200000))).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, -0.0, 1.0, -1.0, 3.0, 1e-9, -7.5e3, 65504.0, 1.17549435e-38 * 8])])
    hi, r, mid, lo = _split(x)
    # residuals exact: recomputed in float64 they agree bit for bit
    assert np.array_equal(r.astype(np.float64), x.astype(np.float64) - hi.astype(np.float64))
    assert np.array_equal(lo.astype(np.float64), r.astype(np.float64) - mid.astype(np.float64))
    # x = hi + mid + lo exactly
    assert np.array_equal(hi.astype(np.float64) + mid.astype(np.float64) + lo.astype(np.float64),
                          x.astype(np.float64))
    # each part is its own high half-word (8 significant bits): BF16 exactly
    for p in (hi, mid, lo):
        assert np.array_equal(_trunc(p), p)
    ax = np.abs(x.astype(np.float64))
    assert np.all(np.abs(mid.astype(np.float64)) <= ax * 2.0 ** -7)
    assert np.all(np.abs(lo.astype(np.float64)) <= ax * 2.0 ** -15)


def test_dropped_products_are_below_fp32_rounding():
    """sum over the three dropped products mid.lo + lo.mid + lo.lo against |a u|."""
    rng = np.random.default_rng(4)
    a = (rng.standard_normal(100000) * np.exp2(rng.integers(-20, 20, 100000))).astype(np.float32)
    u = (rng.standard_normal(100000) * np.exp2(rng.integers(-20, 20, 100000))).astype(np.float32)
    ha, _, ma, la = (p.astype(np.float64) for p in _split(a))
    hu, _, mu, lu = (p.astype(np.float64) for p in _split(u))
    kept = ha * hu + ha * mu + ma * hu + ha * lu + la * hu + ma * mu
    exact = a.astype(np.float64) * u.astype(np.float64)
    rel = np.abs(exact - kept) / np.where(exact != 0, np.abs(exact), 1.0)
    assert rel.max() <= 3 * 2.0 ** -22
