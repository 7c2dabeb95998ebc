"""Parity of the HIP path (through the C-ABI) with the reference.

Goldens are the reference Fortran's own outputs; fresh seeded configurations
are checked against the CPU oracle.  Tolerance: 1e-10 relative (FP64 path)."""
import os
import shutil

import numpy as np
import pytest

from conftest import DECKS, compare_eval, load_deck, load_golden, pkg

pytestmark = pytest.mark.gpu

RTOL = 1e-10
EVAL_FIXTURES = ["eval_n26_t11.npz", "eval_n26_t23.npz", "eval_n26_t44.npz",
                 "eval_n26g_t11.npz", "eval_n26p_t11.npz", "eval_free26_t11.npz",
                 "eval_g58_t11.npz", "eval_g122_t11.npz", "eval_g61p_t11.npz",
                 "eval_n26a_t11.npz", "eval_c26_t11.npz", "eval_g58r1_t11.npz",
                 "eval_g58r10_t11.npz"]

_ctx_cache = {}


def ctx_for(name, twist=(1, 1)):
    key = (name, tuple(twist))
    if key not in _ctx_cache:
        hip = pkg("hip")
        _ctx_cache[key] = hip.Context(load_deck(name), twist=tuple(twist))
    return _ctx_cache[key]


@pytest.mark.parametrize("fn", EVAL_FIXTURES)
def test_hip_matches_reference_goldens(fn):
    print(f"\n[gpu] golden {fn}", flush=True)
    gold = load_golden(fn)
    c = ctx_for(fn.split("_")[1], tuple(gold["twist"]))
    if "delta" in gold and float(gold["delta"]) > 0:
        c.set_delta(float(gold["delta"]))  # the drift cutoff's time step
    out = c.eval(gold["x"])
    errs = compare_eval(out, gold, gold["scalars"].shape[1], RTOL, fn)
    print("   max err", max(errs.values()), flush=True)


@pytest.mark.parametrize("name,nconf", [("n26", 64), ("g58", 64), ("g122", 32), ("g61p", 16)])
def test_hip_matches_oracle_seeded(name, nconf, oracle_mod):
    """Jittered-grid configurations (no coincident electrons, like VMC
    samples); the uniform-random ones are covered by the goldens."""
    print(f"\n[gpu] oracle {name} x{nconf}", flush=True)
    d = load_deck(name)
    dg = pkg("deckgen")
    x = dg.lattice_configs(nconf, d.nptot, d.el[0], 1000 + nconf)
    o = oracle_mod.Oracle(d, twist=(1, 1))
    ref = o.eval(x)
    gold = dict(x=x, scalars=ref["props"][:, :o.info["n_scalar_props"]], g=ref["g"],
                h=ref["h"][:, :d.ntypes], grad2_ph=ref["grad2_ph"],
                rhok=ref["props"][:, o.info["n_scalar_props"]:])
    out = ctx_for(name).eval(x)
    errs = compare_eval(out, gold, o.info["n_scalar_props"], RTOL, name)
    print("   max err", max(errs.values()), flush=True)
    assert np.array_equal(out["status"], ref["status"])


def test_batch_composition_does_not_change_results():
    """A walker's result is the same alone, in a small batch and in a large
    one (walkers share no state)."""
    print("\n[gpu] batch composition", flush=True)
    d = load_deck("g58")
    x = pkg("deckgen").random_configs(40, d.nptot, d.el[0], 5)
    c = ctx_for("g58")
    full = c.eval(x)
    part = c.eval(x[7:9])
    one = c.eval(x[23:24])
    assert np.array_equal(part["props"], full["props"][7:9])
    assert np.array_equal(part["g"], full["g"][7:9])
    assert np.array_equal(one["props"], full["props"][23:24])


def test_underflow_guard_matches_reference_semantics(tmp_path, oracle_mod):
    """With -log psi above ~372 the reference skips the determinant's drift
    and laplacian and leaves grad2_ph as it was (reptation.f:10087-10092)."""
    print("\n[gpu] underflow guard", flush=True)
    src = os.path.join(DECKS, "n26g")
    dst = str(tmp_path / "big")
    shutil.copytree(src, dst, copy_function=shutil.copyfile)
    deck = pkg("deck")
    dg = pkg("deckgen")
    t = deck.read_table(os.path.join(src, "n26.u2"))
    rows = t.coef * 80.0
    dg.write_table(os.path.join(dst, "n26.u2"), rows, t.drt, "scaled")
    d = deck.read_deck(dst)
    x = dg.random_configs(8, d.nptot, d.el[0], 3)
    o = oracle_mod.Oracle(d)
    g2in = np.full(8, 123.25)
    ref = o.eval(x, grad2_in=g2in)
    c = pkg("hip").Context(d)
    out = c.eval(x, grad2_in=g2in)
    assert np.all(ref["status"] == 1)
    assert np.array_equal(out["status"], ref["status"])
    assert np.all(out["grad2_ph"] == 123.25)
    n = o.info["n_scalar_props"]
    gold = dict(scalars=ref["props"][:, :n], g=ref["g"], h=ref["h"], grad2_ph=ref["grad2_ph"],
                rhok=ref["props"][:, n:])
    compare_eval(out, gold, n, RTOL, "guard")


def test_device_pointer_entry_point_matches_host_entry_point():
    print("\n[gpu] device-pointer entry", flush=True)
    import torch
    d = load_deck("g58")
    x = pkg("deckgen").random_configs(16, d.nptot, d.el[0], 9)
    c = ctx_for("g58")
    host = c.eval(x)
    dev = torch.device("cuda:0")
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    props = torch.zeros((16, c.info["n_props"]), dtype=torch.float64, device=dev)
    g = torch.zeros((16, d.nptot, 2), dtype=torch.float64, device=dev)
    h = torch.zeros((16, 2), dtype=torch.float64, device=dev)
    g2 = torch.zeros(16, dtype=torch.float64, device=dev)
    st = torch.zeros(16, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream()
    c.eval_device(xt, props, g, h, g2, st, stream=s)
    torch.cuda.synchronize()
    assert np.array_equal(props.cpu().numpy()[:, 3:], host["props"][:, 3:])
    assert np.array_equal(g.cpu().numpy(), host["g"])
    assert np.array_equal(g2.cpu().numpy(), host["grad2_ph"])


def test_rejects_unsupported_deck_before_any_launch():
    hip = pkg("hip")
    d = load_deck("n26g")
    d.iu3[1] = 0  # up-down pair without u3 while same-spin pairs have it
    d.iu3[2] = 0
    with pytest.raises(hip.QmcError):
        hip.Context(d)


@pytest.mark.parametrize("fn", ["eval_g122_t11.npz", "eval_g58_t11.npz"])
def test_autotuned_phase_chunk_keeps_parity(fn):
    """qmc_autotune picks the pair kernels' phase chunk by timing; the
    chosen chunk is one of its candidates and the context still matches the
    reference's goldens to 1e-10 (the chunk changes the order in which the
    rho_k phases are added, rounding only)."""
    print(f"\n[gpu] autotuned chunk, golden {fn}", flush=True)
    hip = pkg("hip")
    gold = load_golden(fn)
    c = hip.Context(load_deck(fn.split("_")[1]), twist=tuple(gold["twist"]))
    (k64, kmix), (ms64, msmix) = c.autotune(256)
    print(f"   kchunk FP64 {k64} ({ms64:.3f} ms), mixed {kmix} ({msmix:.3f} ms)", flush=True)
    assert k64 >= 1 and kmix >= 1 and ms64 > 0 and msmix > 0
    out = c.eval(gold["x"])
    errs = compare_eval(out, gold, gold["scalars"].shape[1], RTOL, fn)
    print("   max err", max(errs.values()), flush=True)
