"""Per-configuration FP64 rounding study of the reference's evaluation.

For one deck it evaluates many configurations three ways on the CPU:
  ref    the reference Fortran itself (oracle/_ref/refeval, mode 1),
  dbl    the C restatement in double (oracle/libqmc_oracle.so),
  exact  the same restatement in long double (oracle/libqmc_oracle_ld.so):
         the reference's formulas evaluated with a 64-bit significand on the
         same double inputs, i.e. their exact value to about 1e-18,
and writes tests/golden/sweep_<deck>.json: for every quantity of
conftest.config_errors and conftest.h_errors, the distribution over
configurations of the relative errors ref-vs-exact (the reference's own
rounding), dbl-vs-exact and dbl-vs-ref.  The GPU parity tests hold the HIP
path to the exact value with the same per-configuration metric.

Configurations: `nunif` uniform in the box (electrons arbitrarily close)
and `nvmc` sampled by the reference's own VMC (refeval mode 2, eight chains
from jittered grids, every 10th move after 200).

Usage: python tests/golden/sweep_parity.py DECK NUNIF NVMC [DELTA]
(needs oracle/_ref; several minutes at N=122 for 10^4 configurations)
"""
import importlib
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
deck = importlib.import_module("2degas_amd.deck")
deckgen = importlib.import_module("2degas_amd.deckgen")
import pyoracle as po  # noqa: E402
from conftest import config_errors, h_errors  # noqa: E402

DECKS = os.path.join(HERE, "decks")
WORK = "/tmp/degas_sweep"
JOBS = min(8, os.cpu_count() or 1)


def _ref_chunk(args):
    name, k, x = args
    raw, _ = po.run_refeval(os.path.join(DECKS, name), os.path.join(WORK, f"{name}_e{k}"), 1, x)
    d = deck.read_deck(os.path.join(DECKS, name))
    return po.parse_eval(raw, d.nptot)


def _orc_chunk(args):
    name, x, ext = args
    d = deck.read_deck(os.path.join(DECKS, name))
    return po.Oracle(d, extended=ext).eval(x)


def _vmc_chain(args):
    name, k, nmoves, delta = args
    d = deck.read_deck(os.path.join(DECKS, name))
    # a start next to a node of the trial function can leave a single
    # walker stuck (every Langevin move rejected): such a chain is started
    # again from the next jittered grid, at most four times
    for attempt in range(4):
        x0 = deckgen.lattice_configs(1, d.nptot, d.el[0], 500 + k + 1000 * attempt)[0]
        seed = list(d.seed)
        seed[3] = (seed[3] + 17 * k + 5 * attempt) % 4096
        raw, _ = po.run_refeval(os.path.join(DECKS, name), os.path.join(WORK, f"{name}_v{k}"), 2,
                                x0[None], delta=delta, nsteps=nmoves, seed=seed)
        r = po.parse_replay(raw, d.nptot, nmoves)
        acc = float(np.mean(r["age"][200:] == 0))
        if acc >= 0.2:
            return r["x_old"][200::10]
    raise RuntimeError(f"VMC chain {k} of {name} accepts {acc:.3f} of its moves at delta "
                       f"{delta} from four starts: lower the time step")


def _merge(parts):
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def _summary(e):
    out = {}
    for k, v in e.items():
        v = np.asarray(v, dtype=np.float64)
        out[k] = {"max": float(np.max(v)), "p99": float(np.percentile(v, 99)),
                  "p50": float(np.median(v)), "over_1e-10": int(np.sum(v > 1e-10))}
    return out


def main(name, nunif, nvmc, delta):
    os.makedirs(WORK, exist_ok=True)
    d = deck.read_deck(os.path.join(DECKS, name))
    xs = [deckgen.random_configs(nunif, d.nptot, d.el[0], 7000)]
    with ProcessPoolExecutor(JOBS) as ex:
        if nvmc > 0:
            per = (nvmc + JOBS - 1) // JOBS
            chains = list(ex.map(_vmc_chain, [(name, k, 200 + 10 * per, delta) for k in range(JOBS)]))
            xs.append(np.concatenate(chains)[:nvmc])
        x = np.concatenate(xs)
        chunks = np.array_split(x, JOBS * 4)
        ref = _merge(list(ex.map(_ref_chunk, [(name, k, c) for k, c in enumerate(chunks)])))
        dbl = _merge(list(ex.map(_orc_chunk, [(name, c, False) for c in chunks])))
        exa = _merge(list(ex.map(_orc_chunk, [(name, c, True) for c in chunks])))
    info = po.Oracle(d).info
    ns = info["n_scalar_props"]
    ref_full = {"props": ref["scalars"], "g": ref["g"],
                "h": np.pad(ref["h"], ((0, 0), (0, 2 - ref["h"].shape[1])))}
    counts = [d.np[t] for t in range(d.ntypes)]
    nuniq = len(np.unique(x[nunif:].reshape(len(x) - nunif, -1), axis=0)) if len(x) > nunif else 0
    res = {"deck": name, "nptot": d.nptot, "n_uniform": nunif, "n_vmc": int(len(x) - nunif),
           "n_vmc_distinct": int(nuniq),
           "vmc_delta": delta, "compiler": open(os.path.join(po.REF_DIR, "COMPILER")).read().strip()}
    for tag, a, b in (("ref_vs_exact", ref_full, exa), ("dbl_vs_exact", dbl, exa),
                      ("dbl_vs_ref", dbl, ref_full)):
        for part, sl in (("uniform", slice(0, nunif)), ("vmc", slice(nunif, len(x)))):
            if sl.stop - sl.start <= 0:
                continue
            aa = {k: np.asarray(v)[sl] for k, v in a.items()}
            bb = {k: np.asarray(v)[sl] for k, v in b.items()}
            aa["props"], bb["props"] = aa["props"][:, :ns], bb["props"][:, :ns]
            e = config_errors(aa, bb, info, d.ntypes)
            e.update(h_errors(aa, bb, counts))
            res[f"{tag}_{part}"] = _summary(e)
    out = os.path.join(HERE, f"sweep_{name}.json")
    with open(out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", out)
    for key in sorted(k for k in res if k.startswith(("ref_vs", "dbl_vs"))):
        worst = max(res[key].items(), key=lambda kv: kv[1]["max"])
        print(f"  {key:24s} worst {worst[0]:7s} max {worst[1]['max']:.2e} "
              f"p99 {worst[1]['p99']:.2e} over 1e-10: {worst[1]['over_1e-10']}")


if __name__ == "__main__":
    a = sys.argv[1:]
    main(a[0], int(a[1]), int(a[2]), float(a[3]) if len(a) > 3 else 0.1)
