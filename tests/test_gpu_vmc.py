"""Device VMC step (move + metropolis_test + averages) against the
reference's recorded move sequence."""
import numpy as np
import pytest

from conftest import load_deck, load_golden, pkg

pytestmark = pytest.mark.gpu


def _replay(name, tag=None, ctx=None):
    hip = pkg("hip")
    abi = pkg("abi")
    gold = load_golden(f"replay_{tag or name}.npz")
    d = load_deck(name)
    if ctx is None:
        ctx = hip.Context(d, twist=(1, 1))
    every = int(gold["every"])
    states = [abi.limbs_to_state(l) for l in gold["ck_seeds"]]
    wk = hip.Walkers(ctx, gold["ck_x"], states)
    wk.start(delta=float(gold["delta"]))
    tp, ta = wk.sweep(every, float(gold["delta"]), trace=True)
    import torch
    torch.cuda.synchronize()
    assert np.all(wk.status() == 0)
    p = tp.cpu().numpy().T.reshape(-1)       # walker-major = checkpoint order
    age = ta.cpu().numpy().T.reshape(-1)
    return gold, p, age


@pytest.mark.parametrize("name", ["n26", "g58"])
def test_vmc_accept_sequence_matches_reference(name):
    """Teacher-forced replay: every 100-move segment of the reference run
    starts from the reference's recorded walker and RNG state; the device
    reproduces every accept/reject decision and every acceptance
    probability."""
    print(f"\n[gpu] vmc replay {name}", flush=True)
    gold, p, age = _replay(name)
    n = len(gold["p"])
    acc_ref = gold["age"][:n] == 0
    acc = age[:n] == 0
    mism = int(np.sum(acc != acc_ref))
    relp = float(np.max(np.abs(p[:n] - gold["p"]) / np.maximum(gold["p"], 1e-300)))
    print(f"   moves {n} mismatches {mism} max rel p {relp:.2e} acc rate {acc.mean():.3f}", flush=True)
    assert mism == 0
    assert relp < 1e-8


def test_vmc_adrift_two_time_steps_match_reference():
    """The drift cutoff (adrift 0.5) reads the current time step: one
    context replays the reference's VMC record at delta 0.05, then its
    record at delta 0.02; both reproduce every decision and p."""
    print("\n[gpu] vmc replay n26a at delta 0.05 then 0.02", flush=True)
    hip = pkg("hip")
    ctx = hip.Context(load_deck("n26a"), twist=(1, 1))
    for tag in ("n26a_d05", "n26a_d02"):
        gold, p, age = _replay("n26a", tag, ctx)
        n = len(gold["p"])
        mism = int(np.sum((age[:n] == 0) != (gold["age"][:n] == 0)))
        relp = float(np.max(np.abs(p[:n] - gold["p"]) / np.maximum(gold["p"], 1e-300)))
        print(f"   {tag}: moves {n} mismatches {mism} max rel p {relp:.2e}", flush=True)
        assert mism == 0
        assert relp < 1e-8


def test_vmc_block_sums_accumulate_old_properties():
    """averages(2): blk accumulates p_old after every step; with 3 steps the
    energy slot of blk equals the sum of the traced old energies."""
    print("\n[gpu] vmc block sums", flush=True)
    import torch
    hip = pkg("hip")
    abi = pkg("abi")
    d = load_deck("n26")
    ctx = hip.Context(d, twist=(1, 1))
    x = pkg("deckgen").lattice_configs(8, d.nptot, d.el[0], 4)
    sts = [abi.limbs_to_state([3535, 3467, 3262, (3099 + 2 * w) % 4096]) for w in range(8)]
    wk = hip.Walkers(ctx, x, sts)
    wk.start()
    e = []
    for _ in range(3):
        wk.sweep(1, 0.05)
        e.append(wk.props[:, 0].clone())
    torch.cuda.synchronize()
    tot = sum(v.cpu().numpy() for v in e)
    assert np.allclose(wk.blk[:, 0].cpu().numpy(), tot, rtol=1e-14, atol=0)
    assert np.all(wk.blk[:, 2].cpu().numpy() <= 3.0)


def test_vmc_other_draw_order_matches_oracle(oracle_mod):
    """bm_order = 1 (the Box-Muller line's two rannyu() calls evaluated in
    the other order): from the reference's recorded walkers and RNG states,
    the device's accept/reject decisions and acceptance probabilities match
    the oracle's replay in the same order."""
    print("\n[gpu] vmc draw order 1 vs oracle", flush=True)
    import torch
    hip = pkg("hip")
    abi = pkg("abi")
    gold = load_golden("replay_n26.npz")
    d = load_deck("n26")
    nw, nsteps, delta = 4, 40, float(gold["delta"])
    x0 = np.asarray(gold["ck_x"][:nw])
    seeds = [np.asarray(s) for s in gold["ck_seeds"][:nw]]
    ctx = hip.Context(d, twist=(1, 1))
    wk = hip.Walkers(ctx, x0, [abi.limbs_to_state(s) for s in seeds])
    wk.start()
    tp, ta = wk.sweep(nsteps, delta, bm_order=1, trace=True)
    torch.cuda.synchronize()
    assert np.all(wk.status() == 0)
    p, age = tp.cpu().numpy(), ta.cpu().numpy()
    o = oracle_mod.Oracle(d, twist=(1, 1))
    mism, worst = 0, 0.0
    for w in range(nw):
        r = o.vmc_replay(x0[w], nsteps, delta, seeds[w], 1)
        mism += int(np.sum((age[:, w] == 0) != (r["age"] == 0)))
        worst = max(worst, float(np.max(np.abs(p[:, w] - r["p"]) / np.maximum(r["p"], 1e-300))))
    print(f"   moves {nw * nsteps} mismatches {mism} max rel p {worst:.2e}", flush=True)
    assert mism == 0
    assert worst < 1e-8


def test_vmc_graph_replays_match_direct_launches():
    """qmc_vmc_sweep without a trace replays the step as a captured hipGraph;
    a context made with QMC_NO_GRAPH launches every kernel directly.  From the
    same walkers and streams, each replayed sweep leaves exactly the state of
    the direct one (positions, drifts, properties, ages, generator states,
    block sums), with other eager work between the replays and after a change
    of time step (a new capture)."""
    print("\n[gpu] vmc graph replays vs direct launches", flush=True)
    import os
    import torch
    hip = pkg("hip")
    abi = pkg("abi")
    dg = pkg("deckgen")
    d = load_deck("c26")
    x = dg.lattice_configs(64, d.nptot, d.el[0], 31)
    sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * k + 1) % 4096]) for k in range(64)]
    ctx_g = hip.Context(d, twist=(1, 1))
    os.environ["QMC_NO_GRAPH"] = "1"
    try:
        ctx_d = hip.Context(d, twist=(1, 1))
    finally:
        del os.environ["QMC_NO_GRAPH"]
    wg, wd = hip.Walkers(ctx_g, x, sts), hip.Walkers(ctx_d, x, sts)
    wg.start(delta=0.2)
    wd.start(delta=0.2)
    other = torch.tensor(dg.random_configs(16, d.nptot, d.el[0], 5), device="cuda:0")
    for rep, delta in enumerate((0.2, 0.2, 0.2, 0.1, 0.1)):
        wg.sweep(7, delta)
        wd.sweep(7, delta)
        torch.cuda.synchronize()
        for name in ("x", "g", "h", "props", "age", "rng", "blk", "grad2"):
            a, b = getattr(wg, name).cpu().numpy(), getattr(wd, name).cpu().numpy()
            assert np.array_equal(a, b), f"replay {rep}: {name} differs"
        # eager work on the graph context between replays
        ctx_g.eval(other.cpu().numpy())
    assert np.all(wg.status() == 0)
    print(f"   5 sweeps of 7 steps identical, acceptance "
          f"{float((wg.age == 0).double().mean()):.2f}", flush=True)
