"""Comparison helpers for GPU-vs-oracle parity tests.

The FP64 path evaluates the oracle's expressions in the oracle's order, so
its parity tests assert bitwise equality of positions, reference
coordinates, velocities and projected fields.  Relative differences are
also reported, per element (|a - b| / |b|; an exact zero in b counts as a
difference only when a differs from it).
"""
import numpy as np


def _rel(a, b):
    d = np.abs(a - b)
    scale = np.abs(b)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(d == 0.0, 0.0, d / scale)
    both_nan = np.isnan(a) & np.isnan(b)
    r = np.where(both_nan, 0.0, r)
    if np.isnan(r).any():
        return float("inf")
    return float(np.max(r)) if r.size else 0.0


def bitwise_equal(a, b):
    """Bit-for-bit equality of two float64 arrays (NaN payloads included)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(np.array_equal(a.view(np.uint64), b.view(np.uint64)))


def compare_particles(g, o):
    """Return a report of the differences between two particle sets
    (both cell-major, ascending id inside a cell)."""
    rep = {"n_gpu": len(g["id"]), "n_oracle": len(o["id"])}
    if rep["n_gpu"] != rep["n_oracle"]:
        rep["same_count"] = False
        return rep
    rep["same_count"] = True
    rep["cell_equal"] = bool(np.array_equal(g["cell"], o["cell"]))
    rep["id_equal"] = bool(np.array_equal(g["id"], o["id"]))
    for k in ("x", "xi", "v"):
        rep[f"{k}_bitwise"] = bitwise_equal(g[k], o[k])
        rep[f"{k}_maxrel"] = _rel(g[k], o[k])
    return rep


def assert_particles_match(g, o, rel=1e-12, where="", bitwise=False):
    """Same count, cells and ids; x, xi, v bitwise equal (bitwise=True) or
    within `rel` relative per element."""
    rep = compare_particles(g, o)
    assert rep["same_count"], f"{where}: particle counts differ {rep}"
    assert rep["cell_equal"], f"{where}: cell ownership differs"
    assert rep["id_equal"], f"{where}: ids differ"
    for k in ("x", "xi", "v"):
        if bitwise:
            assert rep[f"{k}_bitwise"], f"{where}: {k} not bitwise equal {rep}"
        else:
            assert rep[f"{k}_maxrel"] <= rel, f"{where}: {k} differs {rep}"
    return rep


def field_maxrel(a, b):
    """Largest relative difference per entry (NaN in both counts as equal)."""
    return _rel(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))
