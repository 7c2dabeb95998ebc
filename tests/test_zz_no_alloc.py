"""qmc_eval_batch_device_noalloc, the allocation-free variant of
qmc_eval_batch_device (workspace in the library's fixed per-GPU scratch
array), through the C-ABI: its results equal the allocating entry point's
bit for bit on two calls in a row, the second with fewer walkers, in both
precisions; a batch above the scratch limit is refused with
QMC_ERR_LIMIT before anything is launched."""
import numpy as np
import pytest

from conftest import load_deck, pkg

pytestmark = pytest.mark.gpu


def _bufs(ctx, nw, N):
    import torch
    dev = torch.device("cuda:0")
    f64 = torch.float64
    return (torch.zeros((nw, ctx.info["n_props"]), dtype=f64, device=dev),
            torch.zeros((nw, N, 2), dtype=f64, device=dev),
            torch.zeros((nw, 2), dtype=f64, device=dev),
            torch.zeros(nw, dtype=f64, device=dev),
            torch.zeros(nw, dtype=torch.int32, device=dev))


@pytest.mark.parametrize("name", ["g58", "g122"])
def test_noalloc_matches_allocating_entry_point(name):
    print(f"\n[gpu] qmc_eval_batch_device_noalloc vs qmc_eval_batch_device, {name}", flush=True)
    import torch
    hip, abi, dg = pkg("hip"), pkg("abi"), pkg("deckgen")
    d = load_deck(name)
    ctx = hip.Context(d, twist=(1, 1))
    N = d.nptot
    stream = torch.cuda.current_stream()
    for prec in (abi.QMC_PREC_FP64, abi.QMC_PREC_MIXED):
        for nw in (96, 40):  # the second call smaller than the first
            x = torch.tensor(dg.lattice_configs(nw, N, d.el[0], 500 + nw), dtype=torch.float64,
                             device="cuda:0")
            a = _bufs(ctx, nw, N)
            b = _bufs(ctx, nw, N)
            ctx.eval_device(x, *a, precision=prec, stream=stream)
            ctx.eval_device(x, *b, precision=prec, stream=stream, noalloc=True)
            torch.cuda.synchronize()
            for ta, tb, what in zip(a, b, ("props", "g", "h", "grad2_ph", "status")):
                va, vb = ta.cpu().numpy(), tb.cpu().numpy()
                assert np.array_equal(va, vb, equal_nan=True), f"{name} prec {prec} nw {nw}: {what}"
            print(f"   precision {prec}, {nw} walkers: identical", flush=True)
    # above the limit: refused on the host
    big = 1 << 20
    x = torch.zeros((1, N, 2), dtype=torch.float64, device="cuda:0")
    with pytest.raises(hip.QmcError) as e:
        lib = hip.lib()
        import ctypes as C
        p = C.c_void_p(x.data_ptr())
        rc = lib.qmc_eval_batch_device_noalloc(ctx.h, big, p, p, p, p, p, None, 0, None)
        if rc:
            raise hip.QmcError(rc, "noalloc")
    assert e.value.code == abi.QMC_ERR_LIMIT
    ctx.close()
