"""Per-configuration FP64 parity of the HIP path at the deck shapes of the
workloads, on 1000 uniform and 1000 VMC-sampled configurations per deck.

Each configuration is evaluated three ways: the HIP library, the oracle in
double (the reference's loop order, its own FP64 rounding) and the oracle
in long double (the exact value of the same formulas for the same double
inputs).  Errors are per configuration (conftest.config_errors, h_errors).

* VMC-sampled configurations (walkers after 40 device VMC steps from
  jittered grids): every quantity of every configuration within 1e-10 of
  the exact value and of the double oracle.
* Uniform configurations put electrons arbitrarily close; there the
  reference's own FP64 result misses the exact value by up to 3e-8 (the
  committed tests/golden/sweep_<deck>.json, made by
  tests/golden/sweep_parity.py against the compiled reference).  Each
  configuration is held to max(1e-10, 20 x the double oracle's own error),
  and over the batch the HIP path's largest error must not exceed twice
  the double oracle's.
"""
import os

import numpy as np
import pytest

from conftest import DECKS, config_errors, h_errors, load_deck, pkg

pytestmark = pytest.mark.gpu

NCONF = 1000
# deck and a VMC time step that moves walkers off the grid they start from
CASES = [("c26", 0.2), ("g58", 0.2), ("g58r1", 0.01), ("g58r10", 1.0), ("g122", 1.0),
         ("g61p", 1.0)]


def _errors(a, b, info, d):
    e = config_errors(a, b, info, d.ntypes)
    e.update(h_errors(a, b, [d.np[t] for t in range(d.ntypes)]))
    return e


@pytest.mark.parametrize("name,delta", CASES)
def test_per_configuration_parity(name, delta, oracle_mod):
    print(f"\n[gpu] per-configuration parity {name}: {NCONF} uniform + {NCONF} VMC", flush=True)
    import torch
    hip = pkg("hip")
    abi = pkg("abi")
    dg = pkg("deckgen")
    d = load_deck(name)
    ctx = hip.Context(d, twist=(1, 1))
    xu = dg.random_configs(NCONF, d.nptot, d.el[0], 9100)
    # VMC samples: NCONF walkers, each its own stream, 40 steps
    x0 = dg.lattice_configs(NCONF, d.nptot, d.el[0], 9200)
    states = [abi.limbs_to_state([3535, 3467, 3262, (2 * k + 1) % 4096]) for k in range(NCONF)]
    wk = hip.Walkers(ctx, x0, states, block_sums=False)
    wk.start(delta=delta)
    _, ta = wk.sweep(40, delta, trace=True)
    torch.cuda.synchronize()
    acc = float((ta[10:] == 0).double().mean().item())
    xv = wk.x.cpu().numpy()
    assert acc > 0.2, f"VMC acceptance {acc:.3f}"
    x = np.concatenate([xu, xv])
    out = ctx.eval(x)
    deckdir = os.path.join(DECKS, name)
    dbl = oracle_mod.eval_parallel(deckdir, x, extended=False)
    exa = oracle_mod.eval_parallel(deckdir, x, extended=True)
    info = ctx.info
    ns = info["n_scalar_props"]
    cut = lambda r, sl: {"props": r["props"][sl, :ns], "g": r["g"][sl], "h": r["h"][sl]}
    report = {}
    for part, sl in (("uniform", slice(0, NCONF)), ("vmc", slice(NCONF, 2 * NCONF))):
        e_he = _errors(cut(out, sl), cut(exa, sl), info, d)
        e_de = _errors(cut(dbl, sl), cut(exa, sl), info, d)
        e_hd = _errors(cut(out, sl), cut(dbl, sl), info, d)
        report[part] = (e_he, e_de, e_hd)
        worst_he = max(e_he.items(), key=lambda kv: np.max(kv[1]))
        worst_de = max(e_de.items(), key=lambda kv: np.max(kv[1]))
        print(f"   {part}: HIP-exact worst {worst_he[0]} {np.max(worst_he[1]):.2e}, "
              f"oracle(double)-exact worst {worst_de[0]} {np.max(worst_de[1]):.2e}, "
              f"HIP-oracle {max(np.max(v) for v in e_hd.values()):.2e}", flush=True)
    e_he, e_de, e_hd = report["vmc"]
    bad = {k: float(np.max(v)) for k, v in e_he.items() if np.max(v) > 1e-10}
    bad.update({"vs-oracle " + k: float(np.max(v)) for k, v in e_hd.items() if np.max(v) > 1e-10})
    assert not bad, f"VMC-sampled configurations above 1e-10: {bad}"
    e_he, e_de, _ = report["uniform"]
    for k in e_he:
        lim = np.maximum(1e-10, 20.0 * e_de[k])
        nbad = int(np.sum(e_he[k] > lim))
        assert nbad == 0, f"uniform {k}: {nbad} configurations beyond 20x the oracle's rounding"
        assert np.max(e_he[k]) <= max(2.0 * np.max(e_de[k]), 1e-10), \
            f"uniform {k}: HIP {np.max(e_he[k]):.2e} vs oracle {np.max(e_de[k]):.2e}"
