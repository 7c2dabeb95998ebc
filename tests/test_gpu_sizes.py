"""Parity with the oracle across particle counts: the matrix-core
determinant from 3 to 64 per spin (ragged and padded tiles, one or two
spins above 32), the round-robin pair schedule with one type, two equal
types and two unequal types, and the LDS LU determinant k_det that
QMC_NO_GJ selects.  Decks come from the generator (rs = 5, 40 Ewald
shells); tolerance 1e-10 relative, as for the reference goldens."""
import os

import numpy as np
import pytest

from conftest import compare_eval, pkg

pytestmark = pytest.mark.gpu

RTOL = 1e-10
# 16, 17, 32, 33, 48 and 49 per spin are the tile edges of the matrix-core
# determinant: k_gjm_part's tile skipping below 49 and the switch to k_gjm
SIZES = [(33, 33), (37, 29), (40, 20), (48, 47), (57, 0), (64, 64), (64, 1), (20, 12),
         (29, 29), (13, 13), (5, 3), (37, 0), (16, 16), (17, 15), (32, 31), (49, 48)]
LU_SIZES = [(20, 12), (29, 29), (37, 29)]


def _sizes_case(nup, ndown, tmp_path, oracle_mod, tag):
    dg, deck, hip = pkg("deckgen"), pkg("deck"), pkg("hip")
    path = os.path.join(str(tmp_path), "deck")
    dg.make_deck(path, "sz", nup, ndown, 5.0, 40)
    d = deck.read_deck(path)
    x = dg.lattice_configs(8, d.nptot, d.el[0], 100 + nup + ndown)
    o = oracle_mod.Oracle(d, twist=(1, 1))
    ref = o.eval(x)
    ns = o.info["n_scalar_props"]
    gold = dict(x=x, scalars=ref["props"][:, :ns], g=ref["g"], h=ref["h"][:, :d.ntypes],
                grad2_ph=ref["grad2_ph"], rhok=ref["props"][:, ns:])
    ctx = hip.Context(d, twist=(1, 1))
    try:
        out = ctx.eval(x)
    finally:
        ctx.close()
    errs = compare_eval(out, gold, ns, RTOL, f"{tag} {nup}+{ndown}")
    print("   max err", max(errs.values()), flush=True)
    assert np.array_equal(out["status"], ref["status"])


@pytest.mark.parametrize("nup,ndown", SIZES)
def test_sizes_match_oracle(nup, ndown, tmp_path, oracle_mod):
    print(f"\n[gpu] sizes {nup}+{ndown}", flush=True)
    _sizes_case(nup, ndown, tmp_path, oracle_mod, "gj")


@pytest.mark.parametrize("nup,ndown", LU_SIZES)
def test_lu_determinant_matches_oracle(nup, ndown, tmp_path, oracle_mod, monkeypatch):
    """k_det (LINPACK LU and inverse in LDS), read at context creation."""
    print(f"\n[gpu] LU determinant {nup}+{ndown}", flush=True)
    monkeypatch.setenv("QMC_NO_GJ", "1")
    _sizes_case(nup, ndown, tmp_path, oracle_mod, "lu")


def test_phases_off_lattice_match_oracle(tmp_path, oracle_mod):
    """k components that are few but not a progression: the factorized rho_k
    keeps one sin/cos per phase instead of runs of products.  Each distinct
    kx and ky value is moved by its own relative amount, so both the oracle
    and the library see the same off-lattice k-vectors."""
    print("\n[gpu] off-lattice k components 37+29", flush=True)
    dg, deck, hip = pkg("deckgen"), pkg("deck"), pkg("hip")
    path = os.path.join(str(tmp_path), "deck")
    dg.make_deck(path, "kp", 37, 29, 5.0, 40)
    d = deck.read_deck(path)
    kv = np.array(d.kvec, dtype=np.float64).reshape(-1, 2)
    for c in range(2):
        vals = np.unique(kv[:, c])
        shift = {v: 1.0 + 1e-3 * np.sin(7.0 * i + c) for i, v in enumerate(vals)}
        kv[:, c] = [v * shift[v] for v in kv[:, c]]
    d.kvec = kv.reshape(np.shape(d.kvec))
    x = dg.lattice_configs(8, d.nptot, d.el[0], 321)
    o = oracle_mod.Oracle(d, twist=(1, 1))
    ref = o.eval(x)
    ns = o.info["n_scalar_props"]
    gold = dict(x=x, scalars=ref["props"][:, :ns], g=ref["g"], h=ref["h"][:, :d.ntypes],
                grad2_ph=ref["grad2_ph"], rhok=ref["props"][:, ns:])
    ctx = hip.Context(d, twist=(1, 1))
    try:
        out = ctx.eval(x)
    finally:
        ctx.close()
    errs = compare_eval(out, gold, ns, RTOL, "off-lattice k")
    print("   max err", max(errs.values()), flush=True)
    assert np.array_equal(out["status"], ref["status"])
