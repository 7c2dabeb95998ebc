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


def test_ewald_total_independent_of_alpha():
    """The erfc split gives the same Coulomb energy for two splitting
    parameters (real-space + k-space + constants)."""
    dg = pkg("deckgen")
    from scipy.special import erfc
    N = 14
    L = math.sqrt(math.pi * N)
    e2 = 0.4
    rng = np.random.default_rng(3)
    x = rng.uniform(-L / 2, L / 2, size=(N, 2))
    kv = dg.klist(L, 400)
    tot = []
    for alpha in (1.3, 1.8):
        vk, v0 = dg.ewald_constants(N, L, e2, alpha, kv)
        d = x[:, None, :] - x[None, :, :]
        d -= L * np.round(d / L)
        r = np.sqrt((d ** 2).sum(-1))
        iu = np.triu_indices(N, 1)
        # minimum image suffices where erfc(alpha r) has decayed by L/2
        er = (e2 * erfc(alpha * r[iu]) / r[iu]).sum()
        rho = np.exp(1j * (x @ kv.T)).sum(0)
        ek = (vk * np.abs(rho) ** 2).sum()
        tot.append(er + ek + v0)
    assert abs(tot[0] - tot[1]) < 1e-9 * abs(tot[0])
