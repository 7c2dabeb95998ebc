"""Write the golden fixtures of the parity tests.

Runs the reference Fortran (oracle/_ref/refeval, built from the reference
sources by oracle/build_ref.sh) in this container and stores its inputs and
outputs as small .npz files beside this script.  The decks the fixtures use
live under tests/golden/decks/:

  n26      the reference's shipped N=26 deck (rs=1, backflow, u2, u3, 4x4 twists)
  n26g     the same deck at the Gamma point (ntwists line dropped: no efp)
  n26p     plain Slater-Jastrow: the backflow table replaced by zeros, Gamma
  free26   free plane waves: zero backflow, no u2/u3, Gamma
  c26      generated N=26 (13+13), rs=5: u2 alone, zero backflow,
           60 k-shells (K=232), tables on 200 points (plain Slater-Jastrow)
  g58      generated N=58 (29+29), rs=5, u2 + u3 (100 points) + the
           Gaussian backflow 0.2 exp(-(r-1)^2), 60 k-shells
  g58r1, g58r10   the same at rs=1 and rs=10
  g122     generated N=122 (61+61), rs=10, u2+u3+backflow (eta_yk),
           200 k-shells
  g61p     generated polarized N=61 (+ an empty second type), rs=10
All generated decks come from the restated generators of the reference
(2degas_amd/egas.py through deckgen.make_deck).
  n26r     n26 at the Gamma point without the rhok properties (reptation
           fixtures: 14 property slots per bead)
  n26a     n26 at the Gamma point with the drift cutoff on (adrift 0.5)
  n26r_gs, n26r_gsn, n26r_wp, n26r_rj0, n26r_ec
           n26r with the propagator keywords of reptation_move: gstorto,
           gstorto + nodalaction, wpiu, rejection 0, ecut with a value (with
           the drift cutoff on, without which ecut changes nothing)

Usage: python tests/golden/make_golden.py [rmc|rmcvar|adrift|dmcbranch]
       (needs oracle/_ref; minutes; "rmc" writes only the reptation
       fixtures, "rmcvar" the propagator variants, "adrift" the drift-cutoff
       ones, "dmcbranch" the DMC branching ones)
"""
import importlib
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
deck = importlib.import_module("2degas_amd.deck")
deckgen = importlib.import_module("2degas_amd.deckgen")
import pyoracle as po  # noqa: E402

DECKS = os.path.join(HERE, "decks")
WORK = "/tmp/degas_golden"


# n26r variants: extra deck lines selecting reptation_move's propagator
# options (reptation.f:149-170, 1904-1919, 1973-2101)
RMC_VARIANTS = {
    "n26r_gs": ["gstorto 1"],
    "n26r_gsn": ["gstorto 1", "nodalaction 1"],
    "n26r_wp": ["wpiu 1"],
    "n26r_rj0": ["rejection 0"],
    "n26r_ec": ["ecut 1 -0.4", "adrift 0.5"],
}


def derive_n26_variants():
    src = os.path.join(DECKS, "n26")
    for name in ("n26g", "n26p", "free26", "n26r", "n26a") + tuple(RMC_VARIANTS):
        dst = os.path.join(DECKS, name)
        shutil.rmtree(dst, ignore_errors=True)
        os.makedirs(dst)
        for fn in ("n26.u2", "n26.u3", "n26.b", "n26.v", "n26.k", "runid", "seed",
                   "n26.up.x.res.0", "n26.down.x.res.0"):
            shutil.copy(os.path.join(src, fn), dst)
        with open(os.path.join(src, "n26.sy")) as fh:
            lines = fh.read().splitlines()
        out = []
        for ln in lines:
            w = ln.split()
            if not w:
                continue
            if w[0] == "ntwists":
                continue
            if name.startswith("n26r") and w[0] == "rhok":
                continue
            if w[0] == "end":
                if name == "n26a":
                    out.append("adrift 0.5")
                out.extend(RMC_VARIANTS.get(name, []))
            if name == "free26" and w[0] in ("u2", "u3"):
                continue
            if name in ("n26p", "free26") and w[0] == "backflow":
                ln = ln.replace("n26.b", "zero.b")
            out.append(ln)
        with open(os.path.join(dst, "n26.sy"), "w") as fh:
            fh.write("\n".join(out) + "\n")
        if name in ("n26p", "free26"):
            t = deck.read_table(os.path.join(src, "n26.b"))
            rows = np.zeros_like(t.coef)
            deckgen.write_table(os.path.join(dst, "zero.b"), rows, t.drt, "zero")


def make_generated():
    mk = deckgen.make_deck
    mk(os.path.join(DECKS, "c26"), "c26", 13, 13, 5.0, 60, u3=False, backflow="zero")
    for name, rs in (("g58", 5.0), ("g58r1", 1.0), ("g58r10", 10.0)):
        mk(os.path.join(DECKS, name), name, 29, 29, rs, 60, backflow="gauss")
    mk(os.path.join(DECKS, "g122"), "g122", 61, 61, 10.0, 200)
    mk(os.path.join(DECKS, "g61p"), "g61p", 61, 0, 10.0, 200)


def eval_fixture(name, nconf, seed, twist=(1, 1), extra=None, delta=0.0):
    d = deck.read_deck(os.path.join(DECKS, name))
    x = deckgen.random_configs(nconf, d.nptot, d.el[0], seed)
    if extra is not None:
        x = np.concatenate([extra, x])
    raw, _ = po.run_refeval(os.path.join(DECKS, name), os.path.join(WORK, name), 1, x, twist=twist,
                            delta=delta)
    r = po.parse_eval(raw, d.nptot)
    out = os.path.join(HERE, f"eval_{name}_t{twist[0]}{twist[1]}.npz")
    np.savez_compressed(out, x=x, scalars=r["scalars"], g=r["g"], h=r["h"],
                        grad2_ph=r["grad2_ph"], rhok=r["rhok"], twist=np.array(twist),
                        delta=delta)
    print("wrote", out, x.shape)


def replay_fixture(name, nsteps, every, delta, x0, tag=None):
    d = deck.read_deck(os.path.join(DECKS, name))
    raw, _ = po.run_refeval(os.path.join(DECKS, name), os.path.join(WORK, name + "_r"), 2,
                            x0[None], twist=(1, 1), delta=delta, nsteps=nsteps)
    r = po.parse_replay(raw, d.nptot, nsteps)
    ck = np.arange(0, nsteps, every)
    out = os.path.join(HERE, f"replay_{tag or name}.npz")
    np.savez_compressed(out, delta=delta, every=every, ck_seeds=r["seeds"][ck],
                        ck_x=r["x_old"][ck], p=r["p"], age=r["age"], ltf_new=r["ltf_new"],
                        e_new=r["e_new"], first_seed=r["seeds"][0])
    print("wrote", out, "accept rate", np.mean(r["age"] == 0))


def rmc_fixture(name, nmoves, every, delta, rmc, beads):
    """Reference reptation run (refeval mode 4): the full reptile state every
    `every` moves and every move's draws, weights and decisions."""
    d = deck.read_deck(os.path.join(DECKS, name))
    prm = dict(rmc, every=every)
    raw, _ = po.run_refeval(os.path.join(DECKS, name), os.path.join(WORK, name + "_rmc"), 4,
                            beads, delta=delta, nsteps=nmoves, rmc=prm, deck=d)
    r = po.parse_rmc(raw, d.nptot, nmoves, every, int(rmc["mdelta"]))
    out = os.path.join(HERE, f"rmc_{name}.npz")
    # the bead signs are all 1 on this path (s_new = 1); they are not stored
    keep = {k: v for k, v in r.items() if k != "ck_s"}
    np.savez_compressed(out, delta=delta, every=every, beads=beads,
                        **{k: np.asarray(v) for k, v in rmc.items()}, **keep)
    print("wrote", out, "reptation accept rate", np.mean(r["age"] == 0),
          "bead accept", np.mean(r["pbead"][r["pbead"] >= 0]))


def random_path(x0, ntau, sigma, seed):
    """ntau beads: x0 followed by a Gaussian random walk of step sigma."""
    rng = np.random.default_rng(seed)
    steps = rng.normal(0.0, sigma, size=(ntau - 1,) + x0.shape)
    return np.concatenate([x0[None], x0[None] + np.cumsum(steps, axis=0)])


def make_rmc():
    derive_n26_variants()
    shipped = deck.read_configs([os.path.join(DECKS, "n26", f) for f in
                                 ("n26.up.x.res.0", "n26.down.x.res.0")], [13, 13])
    rmc_fixture("n26r", 3000, 100, 0.05,
                dict(ntau=16, ntauskip=3, mdelta=4, etrial=-0.45, mage=1e9, mage_r=1e9),
                random_path(shipped[0], 16, 0.02, 31))
    make_rmc_variants(shipped[0])
    L58 = deck.read_deck(os.path.join(DECKS, "g58")).el[0]
    g = (np.arange(8) + 0.5) / 8 * L58 - L58 / 2
    sites = np.array([(a, b) for a in g for b in g])
    rng = np.random.default_rng(33)
    x58 = sites[rng.permutation(64)[:58]] + rng.normal(0, 0.05 * L58 / 8, size=(58, 2))
    rmc_fixture("g58", 500, 100, 1.0,
                dict(ntau=10, ntauskip=2, mdelta=3, etrial=-0.3, mage=4, mage_r=6),
                random_path(x58, 10, 0.1, 34))


def make_rmc_variants(x0):
    """One 600-move reptation record per propagator option of the deck."""
    derive_n26_variants()
    for k, name in enumerate(RMC_VARIANTS):
        rmc_fixture(name, 600, 100, 0.01,
                    dict(ntau=16, ntauskip=3, mdelta=4, etrial=-0.45, mage=1e9, mage_r=1e9),
                    random_path(x0, 16, 0.005, 40 + k))


def make_adrift():
    """The drift cutoff: an evaluation at delta 0.05 and two VMC records at
    different time steps on the same deck."""
    derive_n26_variants()
    shipped = deck.read_configs([os.path.join(DECKS, "n26", f) for f in
                                 ("n26.up.x.res.0", "n26.down.x.res.0")], [13, 13])
    eval_fixture("n26a", 16, 22, extra=shipped, delta=0.05)
    replay_fixture("n26a", 1000, 100, 0.05, shipped[0], tag="n26a_d05")
    replay_fixture("n26a", 1000, 100, 0.02, shipped[0], tag="n26a_d02")


def make_dmc_branch():
    """branch2 of the reference (refeval mode 5, one rank) on weight sets:
    populations 1 to 1000, log-normal weights of three spreads, weights
    above the bound 10, equal weights (nothing to comb) and a few very
    small ones; two seeds."""
    rng = np.random.default_rng(31)
    cases = []
    for n in (1, 2, 7, 20, 64, 257, 1000):
        for sd in (0.1, 0.6, 1.5):
            cases.append(rng.lognormal(0.0, sd, size=n))
    cases.append(np.full(50, 0.7))
    w = rng.lognormal(0.0, 0.3, size=100)
    w[[3, 50, 99]] = [12.0, 40.0, 10.5]
    cases.append(w)
    w = rng.lognormal(0.0, 0.3, size=40)
    w[::7] = 1e-9
    cases.append(w)
    seeds = ([3535, 3467, 3262, 1549], [11, 2222, 33, 4001])
    out = {}
    for c, w in enumerate(cases):
        seed = seeds[c % 2]
        raw, _ = po.run_refeval(os.path.join(DECKS, "n26"), os.path.join(WORK, "br"), 5, w, seed=seed)
        r = po.parse_branch(raw, len(w))
        out[f"w_{c}"] = w
        out[f"seed_{c}"] = np.array(seed)
        out[f"src_{c}"] = r["src"]
        out[f"nmult_{c}"] = np.array(r["nmult"])
        out[f"limbs2_{c}"] = np.array(r["limbs2"])
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "dmc_branch.npz"), **out)


def main():
    if not po.have_ref():
        sys.exit("oracle/_ref missing: run oracle/build_ref.sh first")
    os.makedirs(WORK, exist_ok=True)
    derive_n26_variants()
    make_generated()
    d26 = deck.read_deck(os.path.join(DECKS, "n26"))
    shipped = deck.read_configs([os.path.join(DECKS, "n26", f) for f in
                                 ("n26.up.x.res.0", "n26.down.x.res.0")], [13, 13])
    shipped1 = deck.read_configs([os.path.join(DECKS, "n26", f) for f in
                                  ("n26.up.x.res.1", "n26.down.x.res.1")], [13, 13])
    shipped = np.concatenate([shipped, shipped1])
    eval_fixture("n26", 30, 11, (1, 1), extra=shipped)
    eval_fixture("n26", 8, 12, (2, 3))
    eval_fixture("n26", 8, 13, (4, 4))
    eval_fixture("n26g", 16, 14)
    eval_fixture("n26p", 16, 15)
    eval_fixture("free26", 8, 16)
    eval_fixture("g58", 16, 17)
    eval_fixture("g122", 6, 18)
    eval_fixture("g61p", 6, 19)
    eval_fixture("c26", 32, 23)
    eval_fixture("g58r1", 16, 24)
    eval_fixture("g58r10", 16, 25)
    # twist lists written by twistg for a few twists (pins the heap sort)
    gv = {}
    for ij in ((1, 1), (2, 3), (4, 4), (3, 1)):
        x = deckgen.random_configs(1, d26.nptot, d26.el[0], 1)
        raw, _ = po.run_refeval(os.path.join(DECKS, "n26"), os.path.join(WORK, "tw"), 1, x, twist=ij)
        gv[f"t{ij[0]}{ij[1]}"] = po.read_gvec_dump(os.path.join(WORK, "tw"))
    raw, _ = po.run_refeval(os.path.join(DECKS, "n26g"), os.path.join(WORK, "tw"), 1, x)
    gv["gamma"] = po.read_gvec_dump(os.path.join(WORK, "tw"))
    np.savez_compressed(os.path.join(HERE, "twist_n26.npz"), **gv)
    replay_fixture("n26", 10000, 100, 0.05, shipped[0])
    # N=58 starts from a jittered 8x8 grid (58 of its 64 sites)
    L58 = deck.read_deck(os.path.join(DECKS, "g58")).el[0]
    g = (np.arange(8) + 0.5) / 8 * L58 - L58 / 2
    sites = np.array([(a, b) for a in g for b in g])
    rng = np.random.default_rng(21)
    x58 = sites[rng.permutation(64)[:58]] + rng.normal(0, 0.05 * L58 / 8, size=(58, 2))
    replay_fixture("g58", 1000, 100, 0.5, x58)


if __name__ == "__main__":
    os.makedirs(WORK, exist_ok=True)
    if sys.argv[1:] == ["rmc"]:
        make_rmc()
    elif sys.argv[1:] == ["adrift"]:
        make_adrift()
    elif sys.argv[1:] == ["dmcbranch"]:
        make_dmc_branch()
    elif sys.argv[1:] == ["rmcvar"]:
        shipped = deck.read_configs([os.path.join(DECKS, "n26", f) for f in
                                     ("n26.up.x.res.0", "n26.down.x.res.0")], [13, 13])
        make_rmc_variants(shipped[0])
    else:
        main()
        make_rmc()
        make_adrift()
        make_dmc_branch()
