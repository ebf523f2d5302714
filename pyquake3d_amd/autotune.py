"""Tile-shape autotuner for the traction products.

The tile shapes that are chosen at run time are the split-K widths:

* ensemble GEMM (``ens_gemm_kernel``): one wave owns a 64-receiver x
  64-member tile and walks ``kc`` source columns, so the grid is
  ceil(n / 64) x ceil(n / kc) tiles and the step reads ceil(n / kc) partial
  sums per (receiver, member, component);
* dense matvec (``matvec_kernel``): a 256-lane workgroup owns 256 x VN
  receivers (VN = 16 bytes / element size) and walks ``kc`` source columns.

A narrow ``kc`` gives more tiles to fill 256 CUs but more partial-sum
traffic; a wide one the reverse, and where the optimum sits depends on n, the
element type and the device.  The tuner measures every candidate on the
device with the object's own resident matrices (``time_product`` includes the
partial-sum pass), keeps the fastest, and remembers the choice per (device
name, CU count, kernel, n, storage, fix_tn) in a JSON cache so later objects
of the same shape start tuned.  The split width changes only the grouping of
the FP32/FP64 sums, never what is summed.

Cache file: ``$PQ3D_TUNE_CACHE`` if set, else
``~/.cache/pyquake3d_amd/tuning.json``.  ``PQ3D_AUTOTUNE=0`` keeps the
built-in widths (the objects' constructors' defaults) wherever the package
would tune on its own.
"""
import json
import math
import os
import tempfile

from . import _lib

_MAT_NAMES = {_lib.MAT_F64: "f64", _lib.MAT_F32: "f32", _lib.MAT_BF16: "bf16"}
_DENSE_KC = (32, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 4096)


def enabled():
    return os.environ.get("PQ3D_AUTOTUNE", "1") != "0"


def cache_path():
    p = os.environ.get("PQ3D_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "pyquake3d_amd", "tuning.json")


def _load():
    try:
        with open(cache_path()) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key, entry):
    """Merge one entry into the cache file (atomic replace; a cache that
    cannot be written is skipped, the tuned value still applies)."""
    d = _load()
    d[key] = entry
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path), prefix=".tuning.")
        with os.fdopen(fd, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _device_key():
    info = _lib.device_info(0)
    return f"{info['name']}|cu{info['cu']}"


def _nq(obj):
    """Traction components the product computes (2 when Bs and Bd are skipped:
    both zero, or fix_tn)."""
    if getattr(obj, "fix_tn", False):
        return 2
    st = getattr(obj, "structure", None)
    return 2 if (st is not None and st[1]) else 3


def _split(ens):
    """The ensemble runs the BF16-split GEMM."""
    p = getattr(ens, "precision", None)
    return bool(p[1]) if p is not None else False


def _kc_step(ens):
    """Widths are multiples of 32 once the BF16 split is requested (whether or
    not it runs), else of 4."""
    p = getattr(ens, "precision", None)
    return 32 if (p is not None and p[0] != "fp32") else 4


def ensemble_key(ens):
    return f"{_device_key()}|ens_gemm|n{ens.n}|B{ens.B}|nq{_nq(ens)}" + ("|bf16x3" if _split(ens) else "")


def dense_key(sim):
    return (f"{_device_key()}|matvec_{_MAT_NAMES[sim.mat_type]}|n{sim.n}|fixtn{int(bool(sim.fix_tn))}"
            f"|nq{_nq(sim)}")


def ensemble_candidates(n, step=4):
    """kc for 1..32 column chunks, rounded up to the MFMA step (4 sources,
    32 for the BF16-split GEMM; 32 chunks bound the partial-sum memory an
    ensemble is budgeted for)."""
    out = []
    for nc in range(1, 33):
        kc = (math.ceil(n / nc) + step - 1) // step * step
        if kc >= step and kc not in out:
            out.append(kc)
    return out


def dense_candidates(n):
    return [kc for kc in _DENSE_KC if kc < n] + [n] if n > 0 else []


def _sweep(obj, candidates, reps, trials, finalists=4, final_trials=5):
    """Time every candidate (best of `trials`), then re-time the `finalists`
    fastest `final_trials` times each, interleaved so that clock and thermal
    drift hit them alike, and keep the lowest median."""
    table = []
    for kc in candidates:
        obj.set_kc(kc)
        ms = min(obj.time_product(reps) for _ in range(trials))
        table.append((int(obj.kc), float(ms)))
    top = [kc for kc, _ in sorted(table, key=lambda t: t[1])[:finalists]]
    runs = {kc: [] for kc in top}
    for _ in range(final_trials):
        for kc in top:
            obj.set_kc(kc)
            runs[kc].append(obj.time_product(reps))
    med = {kc: sorted(v)[len(v) // 2] for kc, v in runs.items()}
    kc = min(med, key=med.get)
    obj.set_kc(kc)
    return (kc, float(med[kc])), table


def tune_ensemble(ens, candidates=None, reps=4, trials=2, use_cache=True):
    """Measure split-K widths for an ``Ensemble`` (its matrices resident),
    apply the fastest and cache it.  Returns {"kc", "ms", "table", "key"}."""
    key = ensemble_key(ens)
    cands = list(candidates) if candidates is not None else ensemble_candidates(ens.n, _kc_step(ens))
    (kc, ms), table = _sweep(ens, cands, reps, trials)
    entry = {"kc": kc, "ms": ms, "candidates": len(table)}
    if use_cache:
        _store(key, entry)
    return {"kc": kc, "ms": ms, "table": table, "key": key}


def tune_dense(sim, candidates=None, reps=5, trials=2, use_cache=True):
    """Same for a ``DenseSim`` with dense matrices (at its current fix_tn)."""
    if not sim.has_dense:
        raise ValueError("tune_dense: the simulation has no dense matrices")
    key = dense_key(sim)
    cands = list(candidates) if candidates is not None else dense_candidates(sim.n)
    (kc, ms), table = _sweep(sim, cands, reps, trials)
    entry = {"kc": kc, "ms": ms, "candidates": len(table)}
    if use_cache:
        _store(key, entry)
    return {"kc": kc, "ms": ms, "table": table, "key": key}


def cached(key):
    e = _load().get(key)
    return int(e["kc"]) if isinstance(e, dict) and "kc" in e else None


def ensure_ensemble(ens, tune=True):
    """Apply the cached split-K width for this shape, tuning once if absent
    (with ``tune=False`` the built-in width stays when nothing is cached, as
    when other work would share the device during the timing).  Returns the
    kc in use."""
    kc = cached(ensemble_key(ens))
    if kc is None:
        return tune_ensemble(ens)["kc"] if tune else ens.kc
    if kc != ens.kc:
        ens.set_kc(kc)
    return ens.kc


def ensure_dense(sim):
    kc = cached(dense_key(sim))
    if kc is None:
        return tune_dense(sim)["kc"]
    if kc != sim.kc:
        sim.set_kc(kc)
    return sim.kc
