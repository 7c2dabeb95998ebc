"""Device reptation moves (qmc_rmc_start / qmc_rmc_sweep) against the
reference's own read_path, reptation_move and reptation_metropolis_test,
recorded by the reference harness (tests/golden/rmc_*.npz)."""
import numpy as np
import pytest

from conftest import load_deck, load_golden, pkg

pytestmark = pytest.mark.gpu


def _setup(name):
    hip = pkg("hip")
    gold = load_golden(f"rmc_{name}.npz")
    d = load_deck(name)
    ctx = hip.Context(d, twist=(1, 1))
    prm = hip.rmc_params(float(gold["delta"]), float(gold["etrial"]), float(gold["mage"]),
                         float(gold["mage_r"]), int(gold["ntauskip"]), int(gold["mdelta"]),
                         d.driver)
    return hip, gold, d, ctx, prm


def _close(a, b, rtol, atol):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    err = np.abs(a - b)
    bad = err > atol + rtol * np.abs(b)
    return int(np.sum(bad)), float(np.max(err / (atol + rtol * np.abs(b)))) if err.size else 0.0


@pytest.mark.parametrize("name", ["n26r", "g58"])
def test_rmc_start_matches_read_path(name):
    """qmc_rmc_start from the reference's initial beads reproduces the
    reptile read_path builds: bead drifts and properties, link weights,
    Langevin probabilities of the links, path averages and ends."""
    print(f"\n[gpu] rmc start {name}", flush=True)
    import torch
    hip, gold, d, ctx, prm = _setup(name)
    abi = pkg("abi")
    ntau, md = int(gold["ntau"]), int(gold["mdelta"])
    rep = hip.Reptiles(ctx, 1, ntau, md, [abi.limbs_to_state(gold["ck_seed"][0])])
    rep.start(gold["beads"][None], prm)
    torch.cuda.synchronize()
    assert np.all(rep.status() == 0)
    assert list(rep.ends.cpu().numpy()[0]) == [int(gold["ck_jfirst"][0]) - 1,
                                               int(gold["ck_jlast"][0]) - 1]
    g = rep.g.cpu().numpy()[0, :ntau]
    p = rep.props.cpu().numpy()[0, :ntau]
    w = rep.w.cpu().numpy()[0, :ntau]
    pp = rep.pp_old.cpu().numpy()[0]
    scale_p = np.maximum(np.max(np.abs(gold["ck_p"][0][:ntau]), axis=0), 1e-300)
    checks = {
        "g": _close(g, gold["ck_g"][0][:ntau], 1e-9, 1e-11),
        "props": _close(p / scale_p, gold["ck_p"][0][:ntau] / scale_p, 1e-9, 1e-11),
        "w": _close(w, gold["ck_w"][0][:ntau], 1e-9, 1e-9),
        "pp_old": _close(pp / scale_p, gold["ck_pp_old"][0] / scale_p, 1e-9, 1e-11),
    }
    for k, (nbad, worst) in checks.items():
        print(f"   {k}: bad {nbad} worst/tol {worst:.2e}", flush=True)
    assert all(v[0] == 0 for v in checks.values())
    assert float(rep.rage.cpu().numpy()[0]) == float(gold["ck_age"][0])


def _replay(name):
    import torch
    hip, gold, d, ctx, prm = _setup(name)
    abi = pkg("abi")
    ntau, md = int(gold["ntau"]), int(gold["mdelta"])
    every = int(gold["every"])
    nck = len(gold["ck_seed"])
    states = [abi.limbs_to_state(s) for s in gold["ck_seed"]]
    rep = hip.Reptiles(ctx, nck, ntau, md, states)
    dev = rep.x.device
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    rep.x.copy_(t(gold["ck_x"]))
    rep.g.copy_(t(gold["ck_g"]))
    rep.props.copy_(t(gold["ck_p"]))
    rep.age.copy_(t(gold["ck_age_stack"]))
    rep.w.copy_(t(gold["ck_w"]))
    ends = np.stack([gold["ck_jfirst"] - 1, gold["ck_jlast"] - 1], axis=1)
    rep.ends.copy_(t(ends, torch.int32))
    rep.rage.copy_(t(gold["ck_age"]))
    rep.pp_old.copy_(t(gold["ck_pp_old"]))
    rep.grad2.copy_(t(gold["ck_grad2_ph"]))
    tr = rep.sweep(every, prm, trace=True)
    torch.cuda.synchronize()
    assert np.all(rep.status() == 0)
    # reptile c ran moves c*every .. c*every + every - 1
    tr = tr.cpu().numpy().transpose(1, 0, 2).reshape(nck * every, -1)
    return gold, rep, tr, md


@pytest.mark.parametrize("name", ["n26r", "g58", "n26r_gs", "n26r_gsn", "n26r_wp", "n26r_rj0",
                                  "n26r_ec"])
def test_rmc_move_sequence_matches_reference(name):
    """Teacher-forced replay: each 100-move segment starts from the
    reference's recorded reptile and RNG state; the device reproduces every
    direction, every number of beads, every Langevin and reptation decision,
    the Langevin probabilities, the path weights and the path averages.
    The n26r_* decks select the propagator options of reptation_move
    (gstorto, gstorto + nodalaction, wpiu, rejection 0, ecut with the drift
    cutoff on)."""
    print(f"\n[gpu] rmc replay {name}", flush=True)
    gold, rep, tr, md = _replay(name)
    n = len(gold["idir"])
    idir, nd = tr[:n, 0].astype(int), tr[:n, 1].astype(int)
    assert np.array_equal(idir, gold["idir"]), "directions differ"
    assert np.array_equal(nd, gold["ndelta"]), "bead counts differ"
    acc = tr[:n, 3] == 0
    acc_ref = gold["age"][:n] == 0
    mism = int(np.sum(acc != acc_ref))
    pb, pb_ref = tr[:n, 8:8 + md], gold["pbead"][:n]
    assert np.array_equal(pb < 0, pb_ref < 0)
    moved = pb >= 0
    bad_p, worst_p = _close(pb[moved], pb_ref[moved], 1e-8, 1e-12)
    bad_w, worst_w = _close(tr[:n, 2], gold["wnew"][:n], 1e-8, 1e-8)
    bad_e, worst_e = _close(tr[:n, 4], gold["e"][:n], 1e-10, 1e-12)
    bad_l, worst_l = _close(tr[:n, 5], gold["ltf"][:n], 1e-10, 1e-12)
    bad_k, worst_k = _close(tr[:n, 6], gold["kin"][:n], 1e-10, 1e-12)
    print(f"   moves {n} reptation mismatches {mism} accept rate {acc.mean():.3f}", flush=True)
    print(f"   bead p bad {bad_p} ({worst_p:.2e}), path weight bad {bad_w} ({worst_w:.2e}), "
          f"E/ltf/kin bad {bad_e}/{bad_l}/{bad_k} ({worst_e:.2e}/{worst_l:.2e}/{worst_k:.2e})",
          flush=True)
    assert mism == 0
    assert bad_p == 0 and bad_w == 0 and bad_e == 0 and bad_l == 0 and bad_k == 0
    # every Langevin decision: after segment c the reptile equals the
    # reference's next checkpoint over the path window (ends, reptile age,
    # bead ages, positions, drifts and properties)
    ntau, nbuf = int(gold["ntau"]), int(gold["nbuf"])
    ends = rep.ends.cpu().numpy()
    rage = rep.rage.cpu().numpy()
    age = rep.age.cpu().numpy()
    x, g, p = rep.x.cpu().numpy(), rep.g.cpu().numpy(), rep.props.cpu().numpy()
    nck = len(gold["ck_seed"])
    bad_state = 0
    for c in range(nck - 1):
        jf, jl = int(gold["ck_jfirst"][c + 1]) - 1, int(gold["ck_jlast"][c + 1]) - 1
        ok = list(ends[c]) == [jf, jl] and rage[c] == gold["ck_age"][c + 1]
        win = [(jf + i) % nbuf for i in range(ntau)]
        ok = ok and np.array_equal(age[c][win], gold["ck_age_stack"][c + 1][win])
        ok = ok and _close(x[c][win], gold["ck_x"][c + 1][win], 1e-9, 1e-11)[0] == 0
        ok = ok and _close(g[c][win], gold["ck_g"][c + 1][win], 1e-8, 1e-10)[0] == 0
        sc = np.maximum(np.max(np.abs(gold["ck_p"][c + 1][win]), axis=0), 1e-300)
        ok = ok and _close(p[c][win] / sc, gold["ck_p"][c + 1][win] / sc, 1e-9, 1e-11)[0] == 0
        bad_state += 0 if ok else 1
    print(f"   segment end states differing from the next checkpoint: {bad_state}", flush=True)
    assert bad_state == 0


def test_rmc_block_sums_accumulate_path_averages():
    """averages(2) after each reptation test adds p_old = p_p_old: the
    energy slot of blk equals the sum of the traced path energies."""
    print("\n[gpu] rmc block sums", flush=True)
    import torch
    hip, gold, d, ctx, prm = _setup("n26r")
    abi = pkg("abi")
    ntau, md = int(gold["ntau"]), int(gold["mdelta"])
    nr = 4
    sts = [abi.limbs_to_state([3535, 3467, 3262, (3099 + 2 * k) % 4096]) for k in range(nr)]
    rep = hip.Reptiles(ctx, nr, ntau, md, sts)
    beads = np.stack([gold["beads"]] * nr)
    rep.start(beads, prm)
    tr = rep.sweep(7, prm, trace=True)
    torch.cuda.synchronize()
    jetot = ctx.info["jetot"] - 1
    blk = rep.blk.cpu().numpy()[:, jetot]
    ref = tr.cpu().numpy()[:, :, 4].sum(axis=0)
    assert np.allclose(blk, ref, rtol=1e-13, atol=0)


def test_rmc_graph_replays_match_direct_launches():
    """qmc_rmc_sweep without a trace replays the move as a captured hipGraph;
    a context made with QMC_NO_GRAPH launches every kernel directly.  From the
    same reptiles and streams, each replayed sweep leaves exactly the state
    of the direct one (beads, drifts, properties, ages, link weights, ends,
    path averages, generator states, block sums), with eager evaluations
    between the sweeps and after a change of time step (a new capture)."""
    print("\n[gpu] rmc graph replays vs direct launches", flush=True)
    import os
    import torch
    hip, gold, d, _, prm = _setup("n26r")
    abi, dg = pkg("abi"), pkg("deckgen")
    ntau, md = int(gold["ntau"]), int(gold["mdelta"])
    nr = 8
    sts = [abi.limbs_to_state([3535, 3467, 3262, (3099 + 2 * k) % 4096]) for k in range(nr)]
    ctx_g = hip.Context(d, twist=(1, 1))
    os.environ["QMC_NO_GRAPH"] = "1"
    try:
        ctx_d = hip.Context(d, twist=(1, 1))
    finally:
        del os.environ["QMC_NO_GRAPH"]
    beads = np.stack([gold["beads"]] * nr)
    rg, rd = hip.Reptiles(ctx_g, nr, ntau, md, sts), hip.Reptiles(ctx_d, nr, ntau, md, sts)
    rg.start(beads, prm)
    rd.start(beads, prm)
    other = dg.random_configs(4, d.nptot, d.el[0], 6)
    prm2 = hip.rmc_params(float(gold["delta"]) * 0.5, float(gold["etrial"]), float(gold["mage"]),
                          float(gold["mage_r"]), int(gold["ntauskip"]), md, d.driver)
    for rep, p in enumerate((prm, prm, prm2, prm2)):
        rg.sweep(5, p)
        rd.sweep(5, p)
        torch.cuda.synchronize()
        for name in ("x", "g", "props", "age", "w", "ends", "rage", "pp_old", "rng", "blk", "grad2"):
            a, b = getattr(rg, name).cpu().numpy(), getattr(rd, name).cpu().numpy()
            assert np.array_equal(a, b), f"replay {rep}: {name} differs"
        ctx_g.eval(other)
    assert np.all(rg.status() == 0)
    print("   4 sweeps of 5 moves identical", flush=True)
