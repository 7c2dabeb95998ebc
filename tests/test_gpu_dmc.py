"""DMC with fixed-population branching on the device (qmc_dmc_branch,
qmc_dmc_sweep): the branching against the reference's own branch2
(tests/golden/dmc_branch.npz, refeval mode 5), and one sweep step against
its parts run one by one (the VMC step, the weights, the oracle's branch2,
the copies)."""
import numpy as np
import pytest

from conftest import load_deck, load_golden, pkg

pytestmark = pytest.mark.gpu


def test_branch_matches_reference_branch2():
    """Every golden weight set (1 to 1000 walkers, bounded weights, equal
    weights, tiny weights): the same origins for every walker, the same
    count of bounded weights and the second generator left in the same
    state as the reference's branch2."""
    print("\n[gpu] dmc branching vs reference branch2", flush=True)
    import torch
    hip, abi, run = pkg("hip"), pkg("abi"), pkg("run")
    g = load_golden("dmc_branch.npz")
    ctx = hip.Context(load_deck("n26g"), twist=(1, 1))
    dev = torch.device("cuda:0")
    n_cases = int(g["ncases"])
    for c in range(n_cases):
        w = np.asarray(g[f"w_{c}"], dtype=np.float64)
        wt = torch.tensor(w, dtype=torch.float64, device=dev)
        limbs2 = run.setrn_limbs([int(v) for v in g[f"seed_{c}"]])
        rng2 = torch.tensor([abi.limbs_to_state(limbs2)], dtype=torch.int64, device=dev)
        src = torch.zeros(len(w), dtype=torch.int32, device=dev)
        counts = torch.zeros(4, dtype=torch.int32, device=dev)
        hip.dmc_branch(ctx, wt, rng2, src, counts)
        torch.cuda.synchronize()
        got = src.cpu().numpy()
        assert np.array_equal(got, g[f"src_{c}"]), f"case {c} (n={len(w)}): origins differ"
        assert int(counts[1].item()) == int(g[f"nmult_{c}"]), f"case {c}: bounded weights"
        assert abi.state_to_limbs(int(rng2.item())) == [int(v) for v in g[f"limbs2_{c}"]], f"case {c}: rannyu2"
    print(f"   {n_cases} weight sets identical", flush=True)
    ctx.close()


def test_sweep_step_equals_its_parts(oracle_mod):
    """One DMC step (mmstep 2) of 128 walkers on N=58 equals, walker by
    walker: two VMC steps from the same streams (qmc_vmc_sweep), the
    weights exp(-npnorm (elocal - etrial) delta) of each move with elocal as
    metropolis_test leaves it, the weighted block sums, the oracle's branch2
    of the step's weights and the copies it asks for (configuration, drift,
    Laplacian and properties; the age stays)."""
    print("\n[gpu] dmc step vs its parts", flush=True)
    import torch
    hip, abi, dg = pkg("hip"), pkg("abi"), pkg("deckgen")
    d = load_deck("g58")
    ctx = hip.Context(d, twist=(1, 1))
    info = ctx.info
    jetot = info["jetot"] - 1
    nw, delta, etrial, mmstep, mmult = 128, 0.1, -0.36, 2, 3
    x0 = dg.lattice_configs(nw, d.nptot, d.el[0], 4242)
    states = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
    limbs2 = [3535, 3467, 3262, 1549]
    a = hip.Walkers(ctx, x0, states)
    a.start(delta=delta)
    a.sweep(5, delta)          # move off the grid first
    torch.cuda.synchronize()
    b = hip.Walkers(ctx, a.x.cpu().numpy(), [int(v) for v in a.rng.cpu().numpy()])
    for name in ("x", "g", "h", "props", "grad2", "age", "rng"):
        getattr(b, name).copy_(getattr(a, name))
    # the step on the device
    dm = hip.Dmc(a, limbs2)
    a.blk.zero_()
    dm.sweep(1, delta, etrial, mmstep=mmstep, mmult=mmult)
    torch.cuda.synchronize()
    # the parts: plain VMC moves, weights on the host
    wate = np.ones(nw)
    blk = np.zeros((nw, info["n_props"]))
    wsum = np.zeros(nw)
    for m in range(mmstep):
        eold = b.props[:, jetot].cpu().numpy().copy()
        b.sweep(1, delta)
        torch.cuda.synchronize()
        p = b.props.cpu().numpy()
        acc = b.age.cpu().numpy() == 0.0
        el = np.where(acc, 0.5 * (p[:, jetot] + eold), eold)
        peso = np.exp(-float(info["npnorm"]) * (el - etrial) * delta)
        peso = np.minimum(peso, float(mmult))
        wate = wate * peso
        blk += p * wate[:, None]
        wsum += wate
    src, nmult, l2, mult = oracle_mod.branch2(wate, limbs2)
    assert np.sum(mult) == nw
    # the step is a real test of the copies: some walkers are replaced
    assert np.sum(src != np.arange(nw)) > 0
    print(f"   multiplicities: {np.bincount(mult)} copies {int(np.sum(src != np.arange(nw)))}", flush=True)
    assert np.array_equal(dm.src.cpu().numpy(), src)
    assert abi.state_to_limbs(int(dm.rng2.item())) == l2
    np.testing.assert_allclose(dm.wsum.cpu().numpy(), wsum, rtol=1e-13)
    np.testing.assert_allclose(a.blk.cpu().numpy(), blk, rtol=1e-12, atol=1e-300)
    xb, gb, pb = b.x.cpu().numpy(), b.g.cpu().numpy(), b.props.cpu().numpy()
    assert np.array_equal(a.x.cpu().numpy(), xb[src])
    assert np.array_equal(a.g.cpu().numpy(), gb[src])
    assert np.array_equal(a.props.cpu().numpy(), pb[src])
    assert np.array_equal(a.age.cpu().numpy(), b.age.cpu().numpy())
    assert np.array_equal(a.rng.cpu().numpy(), b.rng.cpu().numpy())
    ctx.close()
