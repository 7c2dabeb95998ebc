"""The CPU oracle against the reference's own outputs and closed forms.

Golden vectors come from the reference Fortran itself (oracle/_ref, via
tests/golden/make_golden.py).  The closed-form checks pin pieces the
goldens alone would not isolate."""
import math

import numpy as np
import pytest

from conftest import compare_eval, load_deck, load_golden, pkg

EVAL_FIXTURES = ["eval_n26_t11.npz", "eval_n26_t23.npz", "eval_n26_t44.npz",
                 "eval_n26g_t11.npz", "eval_n26p_t11.npz", "eval_free26_t11.npz",
                 "eval_g58_t11.npz", "eval_g122_t11.npz", "eval_g61p_t11.npz",
                 "eval_n26a_t11.npz", "eval_c26_t11.npz", "eval_g58r1_t11.npz",
                 "eval_g58r10_t11.npz"]


def _deck_of(fn):
    return fn.split("_")[1]


@pytest.mark.parametrize("fn", EVAL_FIXTURES)
def test_oracle_matches_reference(fn, oracle_mod):
    gold = load_golden(fn)
    d = load_deck(_deck_of(fn))
    o = oracle_mod.Oracle(d, twist=tuple(gold["twist"]))
    if "delta" in gold and float(gold["delta"]) > 0:
        o.set_delta(float(gold["delta"]))
    out = o.eval(gold["x"])
    compare_eval(out, gold, gold["scalars"].shape[1], 1e-10, fn)


def test_twist_lists_match_twistg(oracle_mod):
    d = load_deck("n26")
    gv = load_golden("twist_n26.npz")
    for key, ref in gv.items():
        if key == "gamma":
            got = oracle_mod.twist_gvec(d.kvec, d.el, 0, 0, 0, 1)
        else:
            got = oracle_mod.twist_gvec(d.kvec, d.el, d.ntheta, int(key[1]), int(key[2]), 1)
        assert np.array_equal(got, ref), key


def test_heap_sort_is_a_sort_with_reference_ties(oracle_mod):
    # the sorted list is ascending in |g|^2 and a permutation of the input
    d = load_deck("n26")
    g = oracle_mod.twist_gvec(d.kvec, d.el, 4, 3, 1, 1)
    n2 = (g ** 2).sum(1)
    assert np.all(np.diff(n2) >= -1e-12)
    k = d.kvec
    theta = np.array([3, 1]) * (2 * math.pi / np.array(d.el) / 4)
    full = np.concatenate([theta[None], k + theta, -k + theta])
    assert sorted(map(tuple, np.round(full, 12))) == sorted(map(tuple, np.round(g, 12)))


def test_free_plane_waves_kinetic_closed_form(oracle_mod):
    """Zero backflow, no Jastrow: the local kinetic energy of the plane-wave
    determinant is hbs2m * sum_occ |g|^2 / N for any configuration; at Gamma
    for n26 (rs=1) it is the free-particle value 1.04100112 Ry of te.os:22."""
    d = load_deck("free26")
    gold = load_golden("eval_free26_t11.npz")
    o = oracle_mod.Oracle(d)
    out = o.eval(gold["x"])
    info = o.info
    gv = o.gvec
    occ = (gv[:13] ** 2).sum()  # 13 lowest per spin, twice
    ke = d.hbs2m[0] * 2 * occ / 26
    # ekin = hbs2m (lap(-log|psi|) - |grad log|psi||^2); with the phase term
    # hbs2m |grad phi|^2 = grad2_ph the sum is the constant above
    kin = out["props"][:, info["jkin0"] - 1] + out["grad2_ph"] / 26
    assert np.allclose(kin, ke, rtol=1e-10, atol=0)
    assert abs(ke - 1.04100112) < 5e-8


def test_rannyu_known_answer(oracle_mod):
    """Seeding as input does it (reptation.f:37-39) gives the limbs the
    reference prints for n26/seed (3535 3467 3262 3099); the 48-bit LCG
    obeys x_{k} = a^k x_0 mod 2^48 with a = 34522712143931."""
    d = load_deck("n26")
    limbs = [int(v) for v in oracle_mod.rank_seed_limbs(d.seed, 0)]
    assert limbs == [3535, 3467, 3262, 3099]
    import ctypes as C
    L = oracle_mod.lib()
    rng = (C.c_int64 * 4)(*[int(v) for v in limbs])
    a = 34522712143931
    x0 = ((limbs[0] * 4096 + limbs[1]) * 4096 + limbs[2]) * 4096 + limbs[3]
    x = x0
    for k in range(1, 200):
        u = L.orc_rannyu(C.byref(rng))
        x = (a * x) % (1 << 48)
        assert u == x / float(1 << 48)


def test_replay_accept_sequence_n26(oracle_mod):
    """10^4 VMC moves of the reference (move + metropolis_test), replayed in
    teacher-forced segments of 100 from the reference's state and RNG limbs:
    identical accept/reject decisions."""
    gold = load_golden("replay_n26.npz")
    d = load_deck("n26")
    o = oracle_mod.Oracle(d, twist=(1, 1))
    every = int(gold["every"])
    mism = 0
    for c in range(gold["ck_x"].shape[0]):
        r = o.vmc_replay(gold["ck_x"][c], every, float(gold["delta"]), gold["ck_seeds"][c], 0)
        s0 = c * every
        mism += int(np.sum((r["age"] == 0) != (gold["age"][s0:s0 + every] == 0)))
        assert np.allclose(r["ltf_new"], gold["ltf_new"][s0:s0 + every], rtol=1e-8, atol=0)
    assert mism == 0


def test_replay_accept_sequence_g58(oracle_mod):
    gold = load_golden("replay_g58.npz")
    d = load_deck("g58")
    o = oracle_mod.Oracle(d, twist=(1, 1))
    every = int(gold["every"])
    mism = 0
    for c in range(gold["ck_x"].shape[0]):
        r = o.vmc_replay(gold["ck_x"][c], every, float(gold["delta"]), gold["ck_seeds"][c], 0)
        s0 = c * every
        mism += int(np.sum((r["age"] == 0) != (gold["age"][s0:s0 + every] == 0)))
    assert mism == 0


def test_lu_logdet_matches_lapack(oracle_mod):
    """-log|psi| of the free plane-wave deck is -log|det A_up| - log|det A_down|,
    which numpy's LAPACK determinant pins independently of LINPACK."""
    d = load_deck("free26")
    gold = load_golden("eval_free26_t11.npz")
    o = oracle_mod.Oracle(d)
    out = o.eval(gold["x"][:4])
    g = o.gvec[:13]
    for w in range(4):
        tot = 0.0
        for sp in range(2):
            q = gold["x"][w, 13 * sp:13 * sp + 13]
            A = np.exp(1j * (g @ q.T))
            tot -= np.linalg.slogdet(A)[1]
        assert abs(out["props"][w, 1] - tot) < 1e-10 * max(1.0, abs(tot))


def test_drift_cutoff_is_active_and_finite(oracle_mod):
    """The adrift deck (n26a) exercises kinetic's cutoff branch
    (reptation.f:2747-2760): gg < 1 for every configuration, finite drift,
    and the cut drift differs from the uncut one of the same deck read
    without adrift (n26g)."""
    gold = load_golden("eval_n26a_t11.npz")
    jgg = 6
    assert np.all(np.isfinite(gold["scalars"])) and np.all(np.isfinite(gold["g"]))
    assert np.all(gold["scalars"][:, jgg] < 1.0)
    o = oracle_mod.Oracle(load_deck("n26g"))
    plain = o.eval(gold["x"])
    assert np.all(np.abs(plain["g"]) >= np.abs(gold["g"]) - 1e-12)
    assert np.max(np.abs(plain["g"] - gold["g"])) > 1e-6


@pytest.mark.parametrize("tag", ["n26a_d05", "n26a_d02"])
def test_oracle_replays_adrift_vmc(tag, oracle_mod):
    """VMC with the drift cutoff at two time steps: the oracle's replay
    from each 100-move checkpoint of the reference run reproduces its
    accept/reject decisions (first 300 moves)."""
    gold = load_golden(f"replay_{tag}.npz")
    o = oracle_mod.Oracle(load_deck("n26a"))
    every, delta = int(gold["every"]), float(gold["delta"])
    mism = 0
    for c in range(3):
        r = o.vmc_replay(gold["ck_x"][c], every, delta, gold["ck_seeds"][c])
        ref_age = gold["age"][c * every:(c + 1) * every]
        mism += int(np.sum((r["age"] == 0) != (ref_age == 0)))
        assert np.allclose(r["p"], gold["p"][c * every:(c + 1) * every], rtol=1e-8, atol=0)
    assert mism == 0


def test_rel_err_rejects_nan():
    """A NaN anywhere fails a comparison unless the caller marks the slot
    as an empty type's 0/0."""
    from conftest import rel_err
    assert rel_err([np.nan], [np.nan]) == float("inf")
    assert rel_err([1.0], [np.nan]) == float("inf")
    assert rel_err([np.nan], [np.nan], nan_ok=True) == 0.0
    assert rel_err([1.0, np.nan], [1.0, 2.0], nan_ok=True) == float("inf")


def test_branch2_restatement_matches_reference_and_combs(oracle_mod):
    """The oracle's branch2 against the reference's (refeval mode 5) on every
    golden weight set: the same origin for every slot, bounded-weight count
    and second-generator state.  And what the comb must give whatever its
    code: it draws one independent uniform in each unit of the running sum
    of the fractional parts (reptation.f:2196-2210, a stratified comb), so
    every multiplicity is the normalized weight's integer part plus the
    number of those units its fractional part was picked in (at most the
    units its stretch of the running sum meets, at most 2); they add up to
    the population; each slot's configuration comes from a slot with
    multiplicity at least 1 (itself when kept)."""
    po = oracle_mod
    run = pkg("run")
    g = load_golden("dmc_branch.npz")
    for c in range(int(g["ncases"])):
        w = np.asarray(g[f"w_{c}"])
        limbs2 = run.setrn_limbs([int(v) for v in g[f"seed_{c}"]])
        src, nmult, l2, mult = po.branch2(w, limbs2)
        assert np.array_equal(src, g[f"src_{c}"]), c
        assert nmult == int(g[f"nmult_{c}"]) and l2 == [int(v) for v in g[f"limbs2_{c}"]], c
        n = len(w)
        wb = np.minimum(w, 10.0)
        ww = n * wb / wb.sum()
        fl = np.floor(ww)
        cum = np.concatenate([[0.0], np.cumsum(ww - fl)])
        # units [i-1, i) of the running sum each stretch [cum_j, cum_j+1) meets
        meets = np.ceil(cum[1:] - 1e-12) - np.floor(cum[:-1] + 1e-12)
        extra = mult - fl
        assert mult.sum() == n
        assert np.all(extra >= 0) and np.all(extra <= np.minimum(meets, 2)), c
        kept = mult >= 1
        assert np.all(src[kept] == np.arange(n)[kept])
        assert np.all(mult[src] >= 1)
        assert np.array_equal(np.bincount(src, minlength=n), mult)
