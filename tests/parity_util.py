"""Comparison helpers for GPU-vs-oracle parity tests."""
import numpy as np


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
        d = np.abs(g[k] - o[k])
        scale = np.maximum(np.abs(o[k]), 1.0)
        rep[f"{k}_bitwise"] = bool(np.array_equal(g[k].view(np.uint64), o[k].view(np.uint64)))
        rep[f"{k}_maxrel"] = float(np.max(d / scale)) if d.size else 0.0
    return rep


def assert_particles_match(g, o, rel=1e-12, where=""):
    rep = compare_particles(g, o)
    assert rep["same_count"], f"{where}: particle counts differ {rep}"
    assert rep["cell_equal"], f"{where}: cell ownership differs"
    assert rep["id_equal"], f"{where}: ids differ"
    for k in ("x", "xi", "v"):
        assert rep[f"{k}_maxrel"] <= rel, f"{where}: {k} differs {rep}"
    return rep


def field_maxrel(a, b):
    d = np.abs(a - b)
    scale = np.maximum(np.abs(b), 1.0)
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.where(both_nan, 0.0, d)
    if np.isnan(d).any():
        return float("inf")
    return float(np.max(d / scale))
