"""The mixed-precision evaluation (QMC_PREC_MIXED) against the FP64 oracle
at N=122 (61+61, rs=10, u2+u3+backflow, 900 k-vectors) on 10^4
configurations: 5000 uniform in the box and 5000 sampled by device VMC
(1000 walkers, five samples 20 steps apart after 40 steps).

Mixed precision evaluates the u2, u3 and pair-potential tables in FP32 and
forms and sums the rho_k phases in FP32 over chunks of particles; the pair
geometry, the backflow coordinates, the determinant and every accumulator
stay FP64.  Errors per configuration, as the FP64 parity tests measure them:
the local energy against |epot| + |ekin| + |efp|, -log psi against itself,
the drift as max_i |g_i - g_i^ref| / max_i |g_i^ref|.

Tolerances: on VMC-sampled configurations the local energy within 1e-6
(the target of the mixed path), -log psi within 1e-8 and the drift within
1e-6; on uniform configurations, where electrons come arbitrarily close and
FP32 pair terms near r = 0 lose more, the local energy within 1e-5.
"""
import os

import numpy as np
import pytest

from conftest import DECKS, load_deck, pkg

pytestmark = pytest.mark.gpu

NU, NV = 5000, 5000


def _eval(ctx, x, precision):
    import torch
    dev = torch.device("cuda:0")
    N = x.shape[1]
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    nw = x.shape[0]
    props = torch.zeros((nw, ctx.info["n_props"]), dtype=torch.float64, device=dev)
    g = torch.zeros((nw, N, 2), dtype=torch.float64, device=dev)
    h = torch.zeros((nw, 2), dtype=torch.float64, device=dev)
    g2 = torch.zeros(nw, dtype=torch.float64, device=dev)
    ctx.eval_device(xt, props, g, h, g2, precision=precision,
                    stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    return {"props": props.cpu().numpy(), "g": g.cpu().numpy()}


def _errs(a, b, info):
    j = lambda k: info[k] - 1
    pa, pb = a["props"], b["props"]
    es = np.abs(pb[:, j("jpot0")]) + np.abs(pb[:, j("jkin0")]) + np.abs(pb[:, j("jefp")])
    el = np.abs(pa[:, j("jetot")] - pb[:, j("jetot")]) / np.maximum(np.abs(pb[:, j("jetot")]), es)
    ltf = np.abs(pa[:, j("jltf")] - pb[:, j("jltf")]) / np.abs(pb[:, j("jltf")])
    ga, gb = a["g"], b["g"]
    dr = np.max(np.abs(ga - gb).reshape(len(ga), -1), axis=1) / \
        np.max(np.abs(gb).reshape(len(gb), -1), axis=1)
    return {"elocal": el, "ltf": ltf, "drift": dr}


def test_mixed_precision_error_vs_oracle(oracle_mod):
    print(f"\n[gpu] mixed precision vs FP64 oracle, g122, {NU} uniform + {NV} VMC", flush=True)
    import torch
    hip, abi, dg = pkg("hip"), pkg("abi"), pkg("deckgen")
    d = load_deck("g122")
    ctx = hip.Context(d, twist=(1, 1))
    xu = dg.random_configs(NU, d.nptot, d.el[0], 9300)
    nw = NV // 5
    x0 = dg.lattice_configs(nw, d.nptot, d.el[0], 9400)
    states = [abi.limbs_to_state([3535, 3467, 3262, (2 * k + 1) % 4096]) for k in range(nw)]
    wk = hip.Walkers(ctx, x0, states, block_sums=False)
    wk.start(delta=1.0)
    wk.sweep(40, 1.0)
    xv = []
    for _ in range(5):
        wk.sweep(20, 1.0)
        torch.cuda.synchronize()
        xv.append(wk.x.cpu().numpy())
    x = np.concatenate([xu] + xv)
    mix = _eval(ctx, x, abi.QMC_PREC_MIXED)
    f64 = _eval(ctx, x, abi.QMC_PREC_FP64)
    ref = oracle_mod.eval_parallel(os.path.join(DECKS, "g122"), x)
    info = ctx.info
    bad = []
    for part, sl, lim in (("uniform", slice(0, NU), {"elocal": 1e-5, "ltf": 1e-8, "drift": 1e-5}),
                          ("vmc", slice(NU, NU + NV), {"elocal": 1e-6, "ltf": 1e-8, "drift": 1e-6})):
        cut = lambda r: {"props": r["props"][sl], "g": r["g"][sl]}
        em = _errs(cut(mix), cut(ref), info)
        e64 = _errs(cut(f64), cut(ref), info)
        for k in em:
            print(f"   {part} {k}: mixed max {np.max(em[k]):.2e} median {np.median(em[k]):.2e} "
                  f"p99 {np.percentile(em[k], 99):.2e} | FP64 max {np.max(e64[k]):.2e}", flush=True)
            if np.max(em[k]) > lim[k]:
                bad.append(f"{part} {k} {np.max(em[k]):.2e} > {lim[k]:.0e}")
    assert not bad, bad
