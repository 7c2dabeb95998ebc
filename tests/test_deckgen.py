"""Deck generator: formats the reference reads, and Ewald consistency."""
import math
import os

import numpy as np

from conftest import pkg


def test_klist_half_space_and_shells():
    dg = pkg("deckgen")
    L = math.sqrt(26 * math.pi)
    k = dg.klist(L, 18)
    # the shipped n26.k holds the first 18 shells: 56 vectors
    assert k.shape == (56, 2)
    n2 = (k ** 2).sum(1)
    assert np.all(np.diff(n2) >= -1e-12)
    # no vector together with its negative, none zero
    s = set(map(tuple, np.round(k, 9)))
    assert all(tuple(np.round(-v, 9)) not in s for v in k)
    assert np.min(n2) > 0


def test_klist_matches_shipped_n26():
    dg = pkg("deckgen")
    deck = pkg("deck")
    from conftest import DECKS
    ref = deck.read_kvec(os.path.join(DECKS, "n26", "n26.k"))
    k = dg.klist(math.sqrt(26 * math.pi), 18)
    # same vectors shell by shell (ties inside a shell follow the rounding
    # of |k|^2 of the build that wrote the file, so only the sets compare)
    assert k.shape == ref.shape
    assert np.allclose(np.sort((k ** 2).sum(1)), np.sort((ref ** 2).sum(1)), atol=1e-7)
    key = lambda a: sorted(map(tuple, np.round(a, 6)))
    assert key(k) == key(ref)


def test_spline_rows_reproduce_values_and_slopes():
    dg = pkg("deckgen")
    drt = 0.05
    r = np.arange(41) * drt
    f = np.exp(-r) * np.cos(r)
    df = -np.exp(-r) * (np.cos(r) + np.sin(r))
    t = dg.spline_rows(f, df, drt, zero_last=False)
    # the cubic of row i evaluated at rm = 1 hits f_{i+1} and f'_{i+1}
    assert np.allclose(t[:-1].sum(1), f[1:], atol=1e-14)
    assert np.allclose((t[:-1, 1] + 2 * t[:-1, 2] + 3 * t[:-1, 3]) / drt, df[1:], atol=1e-12)


def test_breakup_energy_matches_ewald_sum():
    """A generated deck's Coulomb energy (optimized breakup: real-space
    table with iexp = 1 plus the k-space coefficients plus v0) equals an
    independent erfc Ewald sum of the same periodic system within the
    breakup's fit accuracy (N = 14, rs = 2)."""
    from scipy.special import erfc
    import tempfile
    dg = pkg("deckgen")
    deck = pkg("deck")
    with tempfile.TemporaryDirectory() as td:
        info = dg.make_deck(td, "t14", 7, 7, 2.0, 30, ngrid=400, u3=False, backflow="zero")
        d = deck.read_deck(td)
    N, L, e2 = 14, info["L"], 2.0 / 2.0
    rng = np.random.default_rng(3)
    x = rng.uniform(-L / 2, L / 2, size=(N, 2))
    # the deck's potential: sum over pairs of v_sr(r)/r from the spline
    # rows (minimum image), sum_k vk |rho_k|^2 over the half space, and v0
    t = [tb for tb in d.tables if tb.vk is not None][0]
    dd = x[:, None, :] - x[None, :, :]
    dd -= L * np.round(dd / L)
    r = np.sqrt((dd ** 2).sum(-1))[np.triu_indices(N, 1)]
    ind = np.minimum((r / t.drt).astype(int), t.coef.shape[0] - 1)
    rm = (r - (r / t.drt).astype(int) * t.drt) / t.drt
    c = t.coef[ind]
    vr = c[:, 0] + rm * (c[:, 1] + rm * (c[:, 2] + rm * c[:, 3]))
    rho = np.exp(1j * (x @ d.kvec.T)).sum(0)
    ebreak = (vr / r).sum() + (t.vk * np.abs(rho) ** 2).sum() + d.v0
    # erfc Ewald sum of the same system (alpha large enough for the
    # minimum image in real space, many shells in k space)
    alpha = 1.6
    kv = dg.klist(L, 400)
    k = np.sqrt((kv ** 2).sum(1))
    rho2 = np.abs(np.exp(1j * (x @ kv.T)).sum(0)) ** 2
    A = L * L
    eew = (e2 * erfc(alpha * r) / r).sum() + (2 * np.pi * e2 / (A * k) * erfc(k / (2 * alpha)) * rho2).sum() \
        - N * alpha * e2 / np.sqrt(np.pi) - N * N * e2 * np.sqrt(np.pi) / (alpha * A)
    assert abs(ebreak - eew) < 2e-5 * N * e2, (ebreak, eew)


def test_pade_jastrow_decreases_to_zero_for_every_size(tmp_path):
    """jastrow="pade" (the served jobs' u2): the table decreases over the
    whole range and vanishes at L/2 for N from 26 to 122 and rs from 1 to
    20.  The n26 parameters do not at N=122, rs=5: their rational part
    turns up before L/2 (a well between electrons)."""
    import numpy as np
    dg = pkg("deckgen")
    deck = pkg("deck")
    for n, (nu, nd) in enumerate(((13, 13), (37, 0), (61, 61))):
        for rs in (1.0, 5.0, 20.0):
            dd = str(tmp_path / f"p{n}_{rs}")
            dg.make_deck(dd, "job", nu, nd, rs, 5, u3=False, backflow="zero", jastrow="pade",
                         param_seed=3, perturb=0.1)
            d = deck.read_deck(dd)
            f = d.tables[d.iu2[0] - 1].coef[:, 0]
            assert np.all(np.diff(f) <= 1e-12), (nu, nd, rs)
            assert abs(f[-1]) < 1e-12 and f[0] > 0
    dd = str(tmp_path / "n26_122")
    dg.make_deck(dd, "job", 61, 61, 5.0, 5, u3=False, backflow="zero")
    d = deck.read_deck(dd)
    f = d.tables[d.iu2[0] - 1].coef[:, 0]
    assert np.any(np.diff(f) > 1e-6) and f.min() < -0.5
