import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DECKS = os.path.join(GOLDEN, "decks")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pkg(name=""):
    return importlib.import_module("2degas_amd" + ("." + name if name else ""))


def load_deck(name):
    return pkg("deck").read_deck(os.path.join(DECKS, name))


def load_golden(fn):
    return dict(np.load(os.path.join(GOLDEN, fn)))


def rel_err(a, b, floor=0.0, nan_ok=False):
    """max |a-b| / max(max |b|, floor) (norm-wise relative error).  A NaN or
    infinite entry fails the comparison, unless nan_ok is set (per-type
    slots of an empty type, 0/0 in the reference too): then a and b must be
    non-finite in the same places and of the same kind."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.isfinite(b)
    if not nan_ok and not (np.all(m) and np.all(np.isfinite(a))):
        return float("inf")
    if not np.array_equal(np.isnan(a), np.isnan(b)) or not np.array_equal(np.isinf(a), np.isinf(b)):
        return float("inf")
    if not np.any(m):
        return 0.0
    den = max(np.max(np.abs(b[m])), floor, 1e-300)
    return float(np.max(np.abs(a[m] - b[m])) / den)


def compare_eval(out, gold, n_scalar, rtol, label=""):
    """Compare an evaluation (oracle or HIP) with a reference dump:
    each scalar slot per walker to rtol relative (norm-wise over walkers),
    drift per walker norm-wise, laplacian, grad2_ph and rhok.
    Slot 6 (e2 = E_L^2) carries twice E_L's relative error and is held to
    2 rtol; h, the per-type laplacian sum, adds large terms of both signs
    when electrons come close and is held to 5 rtol (the kinetic energy
    built from it is held to rtol)."""
    sc = out["props"][:, :n_scalar]
    gs = gold["scalars"]
    # the fixed-phase term is rounding noise where the orbitals are real
    # (periodic twists); measure it against the kinetic-energy scale
    ekin_scale = float(np.max(np.abs(gs[:, 4])))
    errs = {}
    for k in range(n_scalar):
        if k == 2:  # acceptance slot is not written by the evaluation
            continue
        floor = ekin_scale if k == n_scalar - 1 else 0.0
        # per-type slots of an empty type are 0/0 or 1/0 in the reference
        nan_ok = not bool(np.any(np.isfinite(gs[:, k])))
        errs[f"slot{k + 1}"] = rel_err(sc[:, k], gs[:, k], floor, nan_ok) / (2.0 if k == 5 else 1.0)
    errs["g"] = max(rel_err(out["g"][w], gold["g"][w]) for w in range(gs.shape[0]))
    nt = gold["h"].shape[1]
    errs["h"] = rel_err(out["h"][:, :nt], gold["h"]) / 5.0
    if np.any(gold["grad2_ph"] != 0):
        npn = out["g"].shape[1]
        errs["grad2_ph"] = rel_err(out["grad2_ph"], gold["grad2_ph"], ekin_scale * npn)
    if gold["rhok"].size:
        errs["rhok"] = rel_err(out["props"][:, n_scalar:], gold["rhok"])
    bad = {k: v for k, v in errs.items() if not v <= rtol}
    assert not bad, f"{label}: errors above {rtol}: {bad}"
    return errs


def config_errors(out, ref, info, ntypes):
    """Per-configuration relative errors of an evaluation `out` against a
    reference evaluation `ref` (dicts with props [nw][n_props], g [nw][N][2],
    h [nw][2]); returns quantity -> array over configurations.

    Scales, per configuration: -log psi and the nodal distances against
    their own magnitude; every energy slot against max(|value|, E_s) with
    E_s = |epot| + |ekin| + |efp| (the terms the local energy adds, the scale
    its rounding acts on); e2 against E_s^2; the drift as max_i |g_i -
    g_i^ref| over max_i |g_i^ref|; the per-type laplacian h_t against
    max(|h_t|, sum over the type of |g_i^ref|^2), the two terms whose
    difference is the kinetic energy."""
    p, q = np.asarray(out["props"]), np.asarray(ref["props"])
    j = lambda k: info[k] - 1

    def rel(a, b, den):
        with np.errstate(invalid="ignore", divide="ignore"):
            r = np.abs(a - b) / den
        return np.where(np.isfinite(a) & np.isfinite(b), r, np.inf)

    es = np.abs(q[:, j("jpot0")]) + np.abs(q[:, j("jkin0")]) + np.abs(q[:, j("jefp")])
    e = {"ltf": rel(p[:, j("jltf")], q[:, j("jltf")], np.abs(q[:, j("jltf")]))}
    for k, name in (("jetot", "elocal"), ("jpot0", "epot"), ("jkin0", "ekin"), ("jefp", "efp")):
        e[name] = rel(p[:, j(k)], q[:, j(k)], np.maximum(np.abs(q[:, j(k)]), es))
    e["e2"] = rel(p[:, j("je2")], q[:, j("je2")], es * es)
    g, gr = np.asarray(out["g"]), np.asarray(ref["g"])
    e["drift"] = np.max(np.abs(g - gr).reshape(len(g), -1), axis=1) / \
        np.max(np.abs(gr).reshape(len(gr), -1), axis=1)
    for t in range(ntypes):
        if info["jkin"][t] == 0:
            continue
        for key, name in (("jkin", "kin"), ("jpot", "pot")):
            k = info[key][t] - 1
            if np.all(~np.isfinite(q[:, k])):
                continue  # an empty type: 0/0 in the reference too
            e[f"{name}{t + 1}"] = rel(p[:, k], q[:, k], np.maximum(np.abs(q[:, k]), es))
        k = info["jun"][t] - 1
        if np.any(np.isfinite(q[:, k])):
            e[f"un{t + 1}"] = rel(p[:, k], q[:, k], np.abs(q[:, k]))
    return e


def per_type_g2(ref, counts):
    """sum over each type of |g_i|^2 of a reference evaluation."""
    g = np.asarray(ref["g"])
    out, o = [], 0
    for n in counts:
        out.append(np.sum(g[:, o:o + n] ** 2, axis=(1, 2)))
        o += n
    return out


def h_errors(out, ref, counts):
    """Per-configuration errors of the per-type laplacian sums against
    max(|h_t|, sum over the type of |g_i^ref|^2)."""
    g2 = per_type_g2(ref, counts)
    e = {}
    for t, n in enumerate(counts):
        if n == 0:
            continue
        a, b = np.asarray(out["h"])[:, t], np.asarray(ref["h"])[:, t]
        e[f"h{t + 1}"] = np.abs(a - b) / np.maximum(np.abs(b), g2[t])
    return e


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.build()
    return pyoracle
