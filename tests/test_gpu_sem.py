"""Single-electron moves with Sherman-Morrison inverse updates
(qmc_sem_refresh / qmc_sem_moves).  The reference has no such move, so the
checks are what the mathematics fixes: the refreshed inverse against
numpy's inverse of the orbital matrix; the inverse carried through a sweep
of rank-one updates against a fresh inverse of the final positions; the
summed log|psi| change of the accepted moves against the difference of two
full evaluations (-ltf, the evaluation the reference's compute_properties
performs); and the VMC energy of the single-electron walk against the
all-electron walk of the reference's move on the same trial function."""
import numpy as np
import pytest

from conftest import load_deck, pkg

pytestmark = pytest.mark.gpu


def _states(nw, base=3099):
    abi = pkg("abi")
    return [abi.limbs_to_state([3535, 3467, 3262, (base + 2 * w) % 4096]) for w in range(nw)]


def _orbital_matrix(gvec, xs):
    # A(k, j) = exp(i g_k . r_j) for the spin's m particles
    m = xs.shape[0]
    return np.exp(1j * (gvec[:m] @ xs.T))


def _inverses(sm):
    a = sm.ainv.cpu().numpy().astype(np.float64)
    return a[..., 0] + 1j * a[..., 1]


def _setup(name, nw, precision=0, deck=None):
    hip = pkg("hip")
    d = deck if deck is not None else load_deck(name)
    ctx = hip.Context(d, twist=(1, 1))
    x = pkg("deckgen").lattice_configs(nw, d.nptot, d.el[0], 7)
    wk = hip.Walkers(ctx, x, _states(nw))
    wk.start()
    sm = hip.SingleMoves(wk, precision=precision)
    return d, ctx, wk, sm


def _spins(d):
    b = [0] + [int(sum(d.np[:t + 1])) for t in range(d.ntypes)]
    return [(b[t], b[t + 1]) for t in range(d.ntypes) if b[t + 1] > b[t]]


def _check_inverse_exact(d, ctx, wk, sm, tol):
    import torch
    torch.cuda.synchronize()
    x = wk.x.cpu().numpy()
    inv = _inverses(sm)
    worst = 0.0
    for w in range(x.shape[0]):
        for sp, (a, b) in enumerate(_spins(d)):
            m = b - a
            A = _orbital_matrix(ctx.gvec, x[w, a:b])
            ref = np.linalg.inv(A)
            worst = max(worst, float(np.max(np.abs(inv[w, sp, :m, :m] - ref)) / np.max(np.abs(ref))))
    assert worst < tol, worst
    return worst


@pytest.mark.parametrize("name", ["c26", "free26"])
def test_refresh_inverse_matches_numpy(name):
    print(f"\n[gpu] sem refresh {name}", flush=True)
    d, ctx, wk, sm = _setup(name, 16)
    sm.refresh()
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-11)
    print(f"   max rel error of the inverse {worst:.2e}", flush=True)


def test_sweeps_keep_inverse_and_logpsi_exact_c26():
    """Five sweeps of single-electron moves: the rank-one-updated inverse
    equals a fresh inverse of the final positions, and the summed
    log|R| - dU of the accepted moves equals -(ltf_after - ltf_before) of
    full evaluations."""
    print("\n[gpu] sem sweeps c26: inverse and log|psi| against full evaluations", flush=True)
    import torch
    d, ctx, wk, sm = _setup("c26", 64)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(5):
        sm.sweep(0.5)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-9)
    wk.start()
    torch.cuda.synchronize()
    dl = sm.dlogpsi.cpu().numpy()
    ref = -(wk.props[:, jltf] - ltf0).cpu().numpy()
    acc = sm.nacc.cpu().numpy() / (5 * d.nptot)
    err = float(np.max(np.abs(dl - ref)))
    print(f"   inverse rel err {worst:.2e}  |dlogpsi - full| {err:.2e}  acceptance {acc.mean():.3f}", flush=True)
    assert 0.2 < acc.mean() < 0.9
    assert err < 1e-8


def test_split_ranges_replay_the_same_walk():
    """The moves of a range draw three uniforms each from the walker's
    stream in order, so moving electrons 0..4, then 5..12, then 13..25
    gives bit for bit the walk of one call over 0..25 (the same streams,
    positions, inverses and accepted counts; the log|psi| sums, grouped per
    call, agree to rounding)."""
    print("\n[gpu] sem split ranges c26", flush=True)
    import torch
    runs = []
    for split in ([0, 26], [0, 5, 13, 26]):
        d, ctx, wk, sm = _setup("c26", 32)
        sm.refresh()
        for _ in range(2):
            for a, b in zip(split[:-1], split[1:]):
                sm.moves(a, b, 0.5)
        torch.cuda.synchronize()
        runs.append([t.cpu().numpy().copy() for t in (wk.x, wk.rng, sm.ainv, sm.dlogpsi, sm.nacc)])
    for k, (u, v) in enumerate(zip(*runs)):
        if k == 3:
            assert np.max(np.abs(u - v)) < 1e-12
        else:
            assert np.array_equal(u, v), k


def _u3_deck(tmp_path, nup=13, ndown=13, rs=5.0, shells=12):
    dg = pkg("deckgen")
    dg.make_deck(str(tmp_path), "u3z", nup, ndown, rs, 5, nshell_k=shells, u3=True, backflow="zero",
                 jastrow="pade")
    return pkg("deck").read_deck(str(tmp_path))


def test_sweeps_with_three_body_jastrow_exact(tmp_path):
    """u2 + u3, zero backflow (generated N=26 deck): the summed
    log|R| - dU of the accepted moves, dU with the three-body term carried
    through the stored per-particle vectors G, equals -(ltf_after -
    ltf_before) of full evaluations, and the inverse stays exact."""
    print("\n[gpu] sem sweeps with u3 (generated N=26 deck)", flush=True)
    import torch
    d = _u3_deck(tmp_path)
    d, ctx, wk, sm = _setup(None, 64, deck=d)
    assert ctx.info["has_det"]
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(5):
        sm.sweep(0.5)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-9)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    acc = float(sm.nacc.sum().item()) / (64 * 5 * d.nptot)
    print(f"   inverse rel err {worst:.2e}  |dlogpsi - full| {err:.2e}  acceptance {acc:.3f}", flush=True)
    assert 0.2 < acc < 0.9
    assert err < 1e-9


def test_energy_with_three_body_agrees_with_all_electron_walk(tmp_path):
    """Statistical agreement with u3 (generated N=26 deck, zero backflow):
    the single-electron walk against the all-electron walk with the drift
    cutoff."""
    print("\n[gpu] sem energy with u3 vs all-electron walk", flush=True)
    d = _u3_deck(tmp_path)
    d.adrift = 0.5
    hip = pkg("hip")
    ctx = hip.Context(d, twist=(1, 1))
    rs2 = 1.0 / d.hbs2m[0]
    ea, nmove = _energy_all_electron(ctx, d, 512, 0.01 * rs2, 400, 2000)
    es, acc = _energy_single(ctx, d, 512, 0.5, 40, 200)
    ma, sa = _blocked(ea)
    ms, ss = _blocked(es)
    z = abs(ma - ms) / np.hypot(sa, ss)
    print(f"   all-electron {ma:.6f} +- {sa:.6f} ({nmove} of 512 walkers move)   single-electron "
          f"{ms:.6f} +- {ss:.6f} (acc {acc:.3f})  z {z:.2f}", flush=True)
    assert nmove == 512
    assert z < 4.5


def test_refresh_tuner_measures_drift():
    """drift() compares the carried inverses with fresh FP64 ones: rounding
    only in FP64, FP32-sized in mixed precision, and growing with the moves
    between refreshes; tune_refresh returns one of its candidates."""
    print("\n[gpu] sem refresh tuner c26", flush=True)
    d, ctx, wk, sm = _setup("c26", 64)
    sm.refresh()
    sm.advance(40, 0.5)
    d64 = sm.drift()
    d, ctx, wk, sm = _setup("c26", 64, precision=1)
    k, drifts = sm.tune_refresh(0.5, candidates=(8, 64, 256), cycles=2)
    print(f"   FP64 drift after 40 moves {d64:.1e}; FP32 drifts {drifts} -> K {k}", flush=True)
    assert d64 < 1e-12
    assert k in (8, 64, 256)
    assert 1e-9 < drifts[8] < 1e-3
    assert drifts[256] < 1e-2


@pytest.mark.parametrize("precision", [0, 1])
def test_packed_jobs_equal_separate_calls(tmp_path, precision):
    """Four jobs of different N, decks and walker counts (c26; a generated
    N=26 deck with u3; N=58 with u3; N=122) moved by packed launches (one per
    spin and row class) give bit for bit what each job's own qmc_sem_moves
    calls give: positions, streams, inverses, log|psi| sums and counts."""
    print(f"\n[gpu] sem packed jobs vs separate calls, precision {precision}", flush=True)
    import torch
    hip = pkg("hip")
    dg = pkg("deckgen")
    decks = [load_deck("c26"), _u3_deck(tmp_path / "a")]
    for tag, nu, sh in (("b", 29, 12), ("c", 61, 12)):
        dg.make_deck(str(tmp_path / tag), "pk", nu, nu, 5.0, 5, nshell_k=sh, u3=tag == "b", backflow="zero",
                     jastrow="pade")
        decks.append(pkg("deck").read_deck(str(tmp_path / tag)))
    sizes, sigmas = [24, 40, 16, 8], [0.5, 0.4, 0.6, 0.5]

    def make():
        out = []
        for k, d in enumerate(decks):
            ctx = hip.Context(d, twist=(1, 1))
            x = dg.lattice_configs(sizes[k], d.nptot, d.el[0], 30 + k)
            wk = hip.Walkers(ctx, x, _states(sizes[k], 500 + 40 * k))
            wk.start()
            sm = hip.SingleMoves(wk, precision=precision)
            sm.refresh()
            out.append(sm)
        return out

    sep, pak = make(), make()
    # tiles other than the defaults on two of the contexts (the 32- and
    # 64-row classes), the same in both runs
    for run in (sep, pak):
        run[2].set_tile(32, 4)
        run[3].set_tile(64, 2)
    for _ in range(2):
        for k, sm in enumerate(sep):
            sm.moves(0, decks[k].nptot, sigmas[k])
        hip.sem_moves_packed([(sm, 0, decks[k].nptot, sigmas[k]) for k, sm in enumerate(pak)], precision)
    torch.cuda.synchronize()
    for k in range(len(decks)):
        a = [t.cpu().numpy() for t in (sep[k].wk.x, sep[k].wk.rng, sep[k].ainv, sep[k].dlogpsi, sep[k].nacc)]
        b = [t.cpu().numpy() for t in (pak[k].wk.x, pak[k].wk.rng, pak[k].ainv, pak[k].dlogpsi, pak[k].nacc)]
        print(f"   job {k} N={decks[k].nptot} walkers {sizes[k]} accepted {a[4].sum():.0f}", flush=True)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("nup,ndown", [(1, 1), (2, 0), (15, 16), (16, 17), (32, 31), (33, 1), (64, 0),
                                       (64, 63)])
def test_sizes_across_row_classes(tmp_path, nup, ndown):
    """Spins of 1 to 64 particles, on both sides of the 16- and 32-row
    class edges and polarized: after two sweeps the updated inverses equal
    fresh ones and the summed log|psi| change equals the full
    evaluations'."""
    print(f"\n[gpu] sem sizes {nup}+{ndown}", flush=True)
    import torch
    dg = pkg("deckgen")
    n = nup + ndown
    shells = 5
    L = float(np.sqrt(n * np.pi))
    while 2 * len(dg.klist(L, shells)) + 1 < max(nup, ndown):
        shells += 2
    dg.make_deck(str(tmp_path), "sz", nup, ndown, 5.0, 5, nshell_k=shells, u3=True, backflow="zero",
                 jastrow="pade")
    d = pkg("deck").read_deck(str(tmp_path))
    d, ctx, wk, sm = _setup(None, 8, deck=d)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(2):
        sm.sweep(0.5)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-8)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    acc = float(sm.nacc.sum().item()) / (8 * 2 * n)
    print(f"   inverse rel err {worst:.1e}  |dlogpsi - full| {err:.1e}  acceptance {acc:.3f}", flush=True)
    assert acc > 0.05
    assert err < 1e-8


@pytest.mark.parametrize("twist", [(2, 3), (4, 1)])
def test_twisted_boundary_conditions(tmp_path, twist):
    """Twisted g-vectors (complex orbitals that are not periodic in the box,
    positions never folded): the updated inverses stay exact and the summed
    log|psi| change equals the full evaluations' at two twists of a 4x4
    grid."""
    print(f"\n[gpu] sem twist {twist}", flush=True)
    import torch
    hip = pkg("hip")
    dg = pkg("deckgen")
    dg.make_deck(str(tmp_path), "tw", 13, 13, 5.0, 5, nshell_k=12, u3=True, backflow="zero", jastrow="pade",
                 ntwists=4)
    d = pkg("deck").read_deck(str(tmp_path))
    assert d.ntheta == 4
    ctx = hip.Context(d, twist=twist)
    x = dg.lattice_configs(16, d.nptot, d.el[0], 5)
    wk = hip.Walkers(ctx, x, _states(16))
    wk.start()
    sm = hip.SingleMoves(wk)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(4):
        sm.sweep(0.7)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-9)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    print(f"   inverse rel err {worst:.1e}  |dlogpsi - full| {err:.1e}", flush=True)
    assert err < 1e-9


@pytest.mark.parametrize("nu,rows", [(29, 32), (61, 64)])
@pytest.mark.parametrize("threads", [2, 4])
def test_tile_shapes_keep_the_walk_exact(tmp_path, nu, rows, threads):
    """Both tile shapes of the 32- and 64-row classes (2 or 4 threads per
    inverse row): the updated inverses stay exact and the log|psi| sums
    equal the full evaluations'; tune_tile picks one of the two and leaves
    the walkers where they were."""
    print(f"\n[gpu] sem tile {rows} rows x {threads} threads", flush=True)
    import torch
    dg = pkg("deckgen")
    dg.make_deck(str(tmp_path), "tl", nu, nu, 5.0, 5, nshell_k=12, u3=True, backflow="zero", jastrow="pade")
    d = pkg("deck").read_deck(str(tmp_path))
    d, ctx, wk, sm = _setup(None, 16, deck=d)
    x0 = wk.x.clone()
    times, best = sm.tune_tile(0.5, sweeps=1)
    torch.cuda.synchronize()
    assert best in (2, 4) and set(times) == {2, 4}
    assert torch.equal(wk.x, x0)
    sm.set_tile(rows, threads)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(2):
        sm.sweep(0.5)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-8)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    print(f"   tuned {times} -> {best}; inverse rel err {worst:.1e}  |dlogpsi - full| {err:.1e}", flush=True)
    assert err < 1e-8


def test_moves_range_checks():
    """Electron ranges out of bounds and a step that is not positive are
    refused."""
    print("\n[gpu] sem argument checks", flush=True)
    hip = pkg("hip")
    d, ctx, wk, sm = _setup("c26", 4)
    with pytest.raises(hip.QmcError):
        sm.moves(0, d.nptot + 1, 0.5)
    with pytest.raises(hip.QmcError):
        sm.moves(3, 2, 0.5)
    with pytest.raises(hip.QmcError):
        sm.moves(0, 4, 0.0)


@pytest.mark.parametrize("name", ["n26g", "g58"])
def test_backflow_moves_track_the_full_evaluation(name):
    """With backflow (u2 + u3 + backflow decks) each move evaluates the
    proposal's log|psi| alone: after three sweeps the ltf kept in the
    walkers' properties equals a fresh full evaluation of their positions,
    and the summed log|psi| change equals -(ltf_after - ltf_before)."""
    print(f"\n[gpu] sem backflow moves {name}", flush=True)
    import torch
    d, ctx, wk, sm = _setup(name, 32)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(3):
        sm.sweep(0.5)
    torch.cuda.synchronize()
    kept = wk.props[:, jltf].cpu().numpy().copy()
    dl = sm.dlogpsi.cpu().numpy()
    acc = float(sm.nacc.sum().item()) / (32 * 3 * d.nptot)
    wk.start()
    torch.cuda.synchronize()
    fresh = wk.props[:, jltf].cpu().numpy()
    kerr = float(np.max(np.abs(kept - fresh)))
    lerr = float(np.max(np.abs(dl + (fresh - ltf0.cpu().numpy()))))
    print(f"   acceptance {acc:.3f}; kept vs fresh ltf {kerr:.1e}; |dlogpsi - full| {lerr:.1e}", flush=True)
    assert 0.1 < acc < 0.95
    assert kerr < 1e-10
    assert lerr < 1e-9


def test_energy_with_backflow_agrees_with_all_electron_walk():
    """Statistical agreement on a backflow deck (n26g: u2 + u3 + the n26
    backflow, Gamma point, rs=1): the single-electron walk, every move
    evaluated in full, against the all-electron walk.  At rs=1 the
    all-electron walk needs the drift cutoff at adrift 1 and a short time
    step (0.005): with adrift 0.5 or delta 0.01, walkers that reach a node
    drift off for thousands of steps (block means jumping by 0.05)."""
    print("\n[gpu] sem energy with backflow vs all-electron walk n26g", flush=True)
    d = load_deck("n26g")
    d.adrift = 1.0
    hip = pkg("hip")
    ctx = hip.Context(d, twist=(1, 1))
    ea, nmove = _energy_all_electron(ctx, d, 512, 0.005, 1000, 4000)
    es, acc = _energy_single(ctx, d, 512, 0.6, 40, 200)
    ma, sa = _blocked(ea)
    ms, ss = _blocked(es)
    z = abs(ma - ms) / np.hypot(sa, ss)
    print(f"   all-electron {ma:.6f} +- {sa:.6f} ({nmove} of 512 walkers move)   single-electron "
          f"{ms:.6f} +- {ss:.6f} (acc {acc:.3f})  z {z:.2f}", flush=True)
    assert nmove >= 500
    assert z < 4.5


def test_mixed_inverse_drift_and_logpsi():
    """QMC_PREC_MIXED keeps the inverse in FP32: after one refresh its
    distance from the FP64 inverse is FP32 rounding; after sweeps of FP32
    rank-one updates it drifts, and a refresh brings it back.  The summed
    log|psi| change stays within 1e-3 of the full evaluations over two
    sweeps."""
    print("\n[gpu] sem mixed precision c26", flush=True)
    import torch
    d, ctx, wk, sm = _setup("c26", 64, precision=1)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    e0 = _check_inverse_exact(d, ctx, wk, sm, 1e-5)
    for _ in range(2):
        sm.sweep(0.5)
    e2 = _check_inverse_exact(d, ctx, wk, sm, 1e-3)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    for _ in range(18):
        sm.sweep(0.5)
    e20 = _check_inverse_exact(d, ctx, wk, sm, 1e-2)
    sm.refresh()
    er = _check_inverse_exact(d, ctx, wk, sm, 1e-5)
    print(f"   inverse rel err: refresh {e0:.1e}, 2 sweeps {e2:.1e}, 20 sweeps {e20:.1e}, "
          f"refreshed {er:.1e}; |dlogpsi - full| after 2 sweeps {err:.1e}", flush=True)
    assert err < 1e-3


def _hard_core_configs(nw, n, L, seed, rmin=0.6):
    """Uniform positions with no two particles closer than rmin (minimum
    image), placed one at a time by rejection."""
    rng = np.random.default_rng(seed)
    x = np.empty((nw, n, 2))
    for w in range(nw):
        k = 0
        while k < n:
            p = rng.uniform(-L / 2, L / 2, 2)
            dd = x[w, :k] - p
            dd -= L * np.round(dd / L)
            if k == 0 or np.min(np.hypot(dd[:, 0], dd[:, 1])) >= rmin:
                x[w, k] = p
                k += 1
    return x


def _energy_all_electron(ctx, d, nw, delta, neq, nmeas):
    """Mean energy per step of the all-electron walk over the walkers that
    move.  Without a drift cutoff, a walker that comes close to a node of
    the determinant gets a drift so large that the drift-diffusion proposal
    is no longer accepted, and stays for thousands of steps (at N=26 and
    delta 0.25, some 8% of walkers within 2000 steps); the deck is therefore
    run with the reference's drift cutoff (adrift 0.5), which bounds the
    drift, and the tests ask that (nearly) every walker moves."""
    import torch
    hip = pkg("hip")
    x = _hard_core_configs(nw, d.nptot, d.el[0], 11)
    wk = hip.Walkers(ctx, x, _states(nw, 1001))
    wk.start(delta=delta)
    wk.sweep(neq, delta)
    je = ctx.info["jetot"] - 1
    es = []
    for _ in range(nmeas):
        wk.sweep(1, delta)
        es.append(wk.props[:, je].clone())
    torch.cuda.synchronize()
    moving = (wk.age < 50).cpu().numpy()
    e = torch.stack(es).cpu().numpy()[:, moving]
    return e.mean(1), int(moving.sum())


def _energy_single(ctx, d, nw, sigma, neq, nmeas, precision=0, refresh_every=10):
    import torch
    hip = pkg("hip")
    x = pkg("deckgen").lattice_configs(nw, d.nptot, d.el[0], 13)
    wk = hip.Walkers(ctx, x, _states(nw, 2001))
    wk.start()
    sm = hip.SingleMoves(wk, precision=precision)
    sm.refresh()
    je = ctx.info["jetot"] - 1
    es = []
    for it in range(neq + nmeas):
        sm.sweep(sigma)
        if (it + 1) % refresh_every == 0:
            sm.refresh()
        if it >= neq:
            wk.start()
            es.append(wk.props[:, je].mean())
    torch.cuda.synchronize()
    acc = float(sm.nacc.sum().item()) / (nw * d.nptot * (neq + nmeas))
    return torch.stack(es).cpu().numpy(), acc


def _blocked(e, nb=20):
    b = np.asarray(e[: len(e) // nb * nb]).reshape(nb, -1).mean(1)
    return float(b.mean()), float(b.std(ddof=1) / np.sqrt(nb))


@pytest.mark.parametrize("precision", [0, 1])
def test_energy_agrees_with_all_electron_walk_c26(precision):
    """Statistical agreement of the VMC energy: the single-electron walk
    (FP64, or with the FP32 inverse refreshed every 10 sweeps) and the
    all-electron drift-diffusion walk (with the drift cutoff adrift 0.5,
    which the single-electron walk does not read) sample the same
    |psi|^2."""
    print(f"\n[gpu] sem energy vs all-electron walk c26 precision {precision}", flush=True)
    d = load_deck("c26")
    d.adrift = 0.5
    hip = pkg("hip")
    ctx = hip.Context(d, twist=(1, 1))
    rs2 = 1.0 / d.hbs2m[0]
    ea, nmove = _energy_all_electron(ctx, d, 512, 0.01 * rs2, 400, 2000)
    es, acc = _energy_single(ctx, d, 512, 0.5, 40, 200, precision=precision)
    ma, sa = _blocked(ea)
    ms, ss = _blocked(es)
    z = abs(ma - ms) / np.hypot(sa, ss)
    print(f"   all-electron {ma:.6f} +- {sa:.6f} ({nmove} of 512 walkers move)   single-electron "
          f"{ms:.6f} +- {ss:.6f} (acc {acc:.3f})  z {z:.2f}", flush=True)
    assert nmove == 512
    assert z < 4.5
    assert np.hypot(sa, ss) < 2e-3 * abs(ma)


def test_sweeps_n122_plain_slater_jastrow(tmp_path):
    """At N=122 (61 per spin, the largest per-spin matrix of the benchmark
    sizes): a generated deck with a Pade two-body Jastrow and no backflow;
    the updated inverse after three sweeps equals a fresh inverse, and the
    summed log|psi| change equals the full evaluations'."""
    print("\n[gpu] sem sweeps N=122 (generated plain Slater-Jastrow deck)", flush=True)
    import torch
    dg = pkg("deckgen")
    dg.make_deck(str(tmp_path), "p122", 61, 61, 5.0, 5, nshell_k=12, u3=False, backflow="zero",
                 jastrow="pade")
    d = pkg("deck").read_deck(str(tmp_path))
    d, ctx, wk, sm = _setup(None, 16, deck=d)
    jltf = ctx.info["jltf"] - 1
    ltf0 = wk.props[:, jltf].clone()
    sm.refresh()
    for _ in range(3):
        sm.sweep(0.5)
    worst = _check_inverse_exact(d, ctx, wk, sm, 1e-8)
    wk.start()
    torch.cuda.synchronize()
    err = float(np.max(np.abs(sm.dlogpsi.cpu().numpy() + (wk.props[:, jltf] - ltf0).cpu().numpy())))
    acc = float(sm.nacc.sum().item()) / (16 * 3 * d.nptot)
    print(f"   inverse rel err {worst:.2e}  |dlogpsi - full| {err:.2e}  acceptance {acc:.3f}", flush=True)
    assert err < 1e-7
