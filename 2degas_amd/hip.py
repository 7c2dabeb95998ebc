"""ctypes binding of lib2degas_hip.so (the HIP walker-evaluation library).

The library is the compute path; this module only marshals arguments.  It
raises when the shared library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib2degas_hip.so")

# every entry point declared in include/qmc_hip.h
EXPORTS = ("qmc_ctx_create", "qmc_ctx_destroy", "qmc_ctx_info", "qmc_set_twist",
           "qmc_twist_gvec", "qmc_eval_batch", "qmc_eval_batch_device", "qmc_strerror",
           "qmc_last_error", "qmc_flops_per_eval", "qmc_timing", "qmc_timing_read",
           "qmc_vmc_start", "qmc_vmc_sweep", "qmc_vmc_status", "qmc_rmc_start",
           "qmc_rmc_sweep", "qmc_set_delta", "qmc_eval_batch_device_noalloc", "qmc_dmc_sweep",
           "qmc_dmc_branch", "qmc_sem_refresh", "qmc_sem_moves",
           "qmc_sem_moves_packed", "qmc_autotune", "qmc_sem_set_tile")

_lib = None


class QmcError(RuntimeError):
    def __init__(self, code, where):
        L = lib()
        msg = L.qmc_last_error().decode() or L.qmc_strerror(code).decode()
        super().__init__(f"{where}: {msg} (code {code})")
        self.code = code


def lib():
    """Load the HIP library (raises if it has not been built).

    torch's wheel bundles a HIP runtime with the same soname as the system
    one; loading torch first makes the library bind to that runtime so the
    process holds one HIP runtime and torch-allocated device buffers can be
    passed to the device-pointer entry points."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        dp, ip, vp = abi.c_double_p, abi.c_int32_p, C.c_void_p
        L.qmc_ctx_create.argtypes = [C.POINTER(abi.QmcTables), C.c_int, C.POINTER(vp)]
        L.qmc_ctx_destroy.argtypes = [vp]
        L.qmc_ctx_info.argtypes = [vp, C.POINTER(abi.QmcInfo)]
        L.qmc_set_twist.argtypes = [vp, dp, C.c_int]
        L.qmc_set_delta.argtypes = [vp, C.c_double]
        L.qmc_twist_gvec.argtypes = [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int, dp]
        L.qmc_eval_batch.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, ip]
        L.qmc_eval_batch_device.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, C.c_int, vp]
        L.qmc_eval_batch_device_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, C.c_int, vp]
        L.qmc_strerror.restype = C.c_char_p
        L.qmc_strerror.argtypes = [C.c_int]
        L.qmc_last_error.restype = C.c_char_p
        L.qmc_flops_per_eval.restype = C.c_double
        L.qmc_flops_per_eval.argtypes = [C.c_int, C.c_int, C.c_int]
        L.qmc_timing.argtypes = [vp, C.c_int]
        L.qmc_vmc_start.argtypes = [vp, C.POINTER(abi.QmcWalkers), vp]
        L.qmc_vmc_sweep.argtypes = [vp, C.POINTER(abi.QmcWalkers), C.c_int, C.c_double,
                                    C.c_double, C.c_int, vp, vp, vp]
        L.qmc_vmc_status.argtypes = [vp, C.c_int, ip]
        L.qmc_rmc_start.argtypes = [vp, C.POINTER(abi.QmcReptiles), C.POINTER(abi.QmcRmcParams), vp]
        L.qmc_rmc_sweep.argtypes = [vp, C.POINTER(abi.QmcReptiles), C.POINTER(abi.QmcRmcParams),
                                    C.c_int, vp, vp]
        L.qmc_timing_read.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.qmc_dmc_sweep.argtypes = [vp, C.POINTER(abi.QmcWalkers), C.POINTER(abi.QmcDmcState), C.c_int,
                                    C.POINTER(abi.QmcDmcParams), vp]
        L.qmc_dmc_branch.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.qmc_sem_refresh.argtypes = [vp, C.POINTER(abi.QmcWalkers), C.POINTER(abi.QmcSemState), C.c_int, vp]
        L.qmc_sem_moves.argtypes = [vp, C.POINTER(abi.QmcWalkers), C.POINTER(abi.QmcSemState), C.c_int,
                                    C.c_int, C.c_double, C.c_int, vp]
        L.qmc_sem_moves_packed.argtypes = [C.POINTER(abi.QmcSemJob), C.c_int, C.c_int, vp]
        L.qmc_autotune.argtypes = [vp, C.c_int, ip, dp, vp]
        L.qmc_sem_set_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def twist_gvec(kvec, el, ntheta, ith, jth):
    """g-vectors of twist (ith, jth) as tabc builds them (host code)."""
    kvec = np.ascontiguousarray(kvec, dtype=np.float64).reshape(-1, 2)
    el = np.ascontiguousarray(el, dtype=np.float64)
    out = np.zeros((2 * kvec.shape[0] + 1, 2))
    rc = lib().qmc_twist_gvec(abi.dptr(kvec), kvec.shape[0], abi.dptr(el), ntheta, ith, jth,
                              abi.dptr(out))
    if rc:
        raise QmcError(rc, "qmc_twist_gvec")
    return out


def flops_per_eval(nup, ndown, nk_ewald):
    return lib().qmc_flops_per_eval(nup, ndown, nk_ewald)


class Context:
    """A device context for one deck (qmc_ctx_create + qmc_set_twist)."""

    def __init__(self, deck, twist=(1, 1), gvec=None, device=0):
        self.deck = deck
        s, keep = deck.pack()
        h = C.c_void_p()
        rc = lib().qmc_ctx_create(C.byref(s), device, C.byref(h))
        if rc:
            raise QmcError(rc, "qmc_ctx_create")
        self.h = h
        if gvec is None:
            gvec = twist_gvec(deck.kvec, deck.el, deck.ntheta, twist[0], twist[1])
        self.set_twist(gvec)
        inf = abi.QmcInfo()
        lib().qmc_ctx_info(self.h, C.byref(inf))
        self.info = inf.as_dict()

    def device_bytes(self):
        """qmc_ctx_info's memory fields now: dev_bytes (what the context
        holds) and the per-walker bytes of its batch-sized buffers."""
        inf = abi.QmcInfo()
        rc = lib().qmc_ctx_info(self.h, C.byref(inf))
        if rc:
            raise QmcError(rc, "qmc_ctx_info")
        d = inf.as_dict()
        return {k: d[k] for k in ("dev_bytes", "eval_bytes_per_walker", "vmc_bytes_per_walker",
                                  "stage_bytes_per_walker")}

    def set_twist(self, gvec):
        self.gvec = np.ascontiguousarray(gvec, dtype=np.float64).reshape(-1, 2)
        rc = lib().qmc_set_twist(self.h, abi.dptr(self.gvec), self.gvec.shape[0])
        if rc:
            raise QmcError(rc, "qmc_set_twist")

    def autotune(self, nw=1024, stream=None):
        """qmc_autotune: the pair kernels' phase chunk picked by timing nw
        synthetic configurations; returns ((kchunk FP64, kchunk mixed),
        (ms FP64, ms mixed))."""
        import torch
        kc = np.zeros(2, np.int32)
        ms = np.zeros(2)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().qmc_autotune(self.h, int(nw), abi.iptr(kc), abi.dptr(ms), C.c_void_p(st.cuda_stream))
        if rc:
            raise QmcError(rc, "qmc_autotune")
        return (int(kc[0]), int(kc[1])), (float(ms[0]), float(ms[1]))

    def set_delta(self, delta):
        """Time step read by the drift cutoff (decks with adrift != 0)."""
        rc = lib().qmc_set_delta(self.h, float(delta))
        if rc:
            raise QmcError(rc, "qmc_set_delta")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().qmc_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eval(self, x, grad2_in=None):
        """Host-pointer batched evaluation (qmc_eval_batch)."""
        N = self.deck.nptot
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, N, 2)
        nw = x.shape[0]
        props = np.zeros((nw, self.info["n_props"]))
        g = np.zeros((nw, N, 2))
        h = np.zeros((nw, 2))
        g2 = np.zeros(nw) if grad2_in is None else np.array(grad2_in, dtype=np.float64)
        st = np.zeros(nw, np.int32)
        rc = lib().qmc_eval_batch(self.h, nw, abi.dptr(x), abi.dptr(props), abi.dptr(g),
                                  abi.dptr(h), abi.dptr(g2), abi.iptr(st))
        if rc:
            raise QmcError(rc, "qmc_eval_batch")
        return dict(props=props, g=g, h=h, grad2_ph=g2, status=st)

    def timing(self, enable):
        """Reset and enable (or disable) per-kernel event timing."""
        rc = lib().qmc_timing(self.h, 1 if enable else 0)
        if rc:
            raise QmcError(rc, "qmc_timing")

    def timing_read(self):
        """Mean ms per launch of (pairs, determinant, drift terms, cross,
        final) and the launch count."""
        ms = np.zeros(5)
        n = C.c_int(0)
        rc = lib().qmc_timing_read(self.h, abi.dptr(ms), C.byref(n))
        if rc:
            raise QmcError(rc, "qmc_timing_read")
        return ms, n.value

    def eval_device(self, x, props, g, h, grad2, status=None, precision=abi.QMC_PREC_FP64,
                    stream=None, noalloc=False):
        """Device-pointer evaluation; arguments are torch tensors on the GPU
        (or raw integer addresses).  Only enqueues on `stream`.  noalloc
        selects qmc_eval_batch_device_noalloc (workspace in the library's
        fixed per-GPU scratch array)."""
        def ptr(t):
            if t is None:
                return None
            return C.c_void_p(t if isinstance(t, int) else t.data_ptr())
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        nw = x.shape[0]
        fn = lib().qmc_eval_batch_device_noalloc if noalloc else lib().qmc_eval_batch_device
        rc = fn(self.h, nw, ptr(x), ptr(props), ptr(g), ptr(h), ptr(grad2), ptr(status), precision, st)
        if rc:
            raise QmcError(rc, "qmc_eval_batch_device" + ("_noalloc" if noalloc else ""))


class Walkers:
    """Device walker state of a batched VMC run (torch tensors on cuda)."""

    @staticmethod
    def state_bytes(nw, nptot, n_props, block_sums=True):
        """Bytes of the tensors __init__ allocates for nw walkers."""
        per = 8 * (2 * nptot * 2 + 2 + n_props + 1 + 1 + 1) + (8 * n_props if block_sums else 0)
        return nw * per

    def __init__(self, ctx, x, rng_states, block_sums=True):
        import torch
        self.ctx = ctx
        dev = torch.device("cuda", 0)
        N = ctx.deck.nptot
        nw = x.shape[0]
        f64 = torch.float64
        self.nw = nw
        self.x = torch.tensor(np.asarray(x, dtype=np.float64).reshape(nw, N, 2), device=dev)
        self.g = torch.zeros((nw, N, 2), dtype=f64, device=dev)
        self.h = torch.zeros((nw, 2), dtype=f64, device=dev)
        self.props = torch.zeros((nw, ctx.info["n_props"]), dtype=f64, device=dev)
        self.grad2 = torch.zeros(nw, dtype=f64, device=dev)
        self.age = torch.zeros(nw, dtype=f64, device=dev)
        self.rng = torch.tensor(np.asarray(rng_states, dtype=np.int64), device=dev)
        self.blk = torch.zeros((nw, ctx.info["n_props"]), dtype=f64, device=dev) if block_sums else None
        s = abi.QmcWalkers()
        s.nw = nw
        for name in ("x", "g", "h", "props", "age", "rng", "blk"):
            t = getattr(self, name)
            setattr(s, name, t.data_ptr() if t is not None else None)
        s.grad2_ph = self.grad2.data_ptr()
        self.s = s

    def _stream(self, stream):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(st.cuda_stream)

    def start(self, stream=None, delta=None):
        """read_conf's start; delta (the vmc record's time step) is what the
        drift cutoff of an adrift deck reads during this evaluation."""
        if delta is not None:
            self.ctx.set_delta(delta)
        rc = lib().qmc_vmc_start(self.ctx.h, C.byref(self.s), self._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_vmc_start")

    def sweep(self, nsteps, delta, mage=1e30, bm_order=0, trace=False, stream=None):
        import torch
        tp = ta = None
        if trace:
            tp = torch.zeros((nsteps, self.nw), dtype=torch.float64, device=self.x.device)
            ta = torch.zeros((nsteps, self.nw), dtype=torch.float64, device=self.x.device)
        rc = lib().qmc_vmc_sweep(self.ctx.h, C.byref(self.s), nsteps, delta, mage, bm_order,
                                 C.c_void_p(tp.data_ptr()) if trace else None,
                                 C.c_void_p(ta.data_ptr()) if trace else None,
                                 self._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_vmc_sweep")
        return tp, ta

    def status(self):
        st = np.zeros(self.nw, np.int32)
        rc = lib().qmc_vmc_status(self.ctx.h, self.nw, abi.iptr(st))
        if rc:
            raise QmcError(rc, "qmc_vmc_status")
        return st


class Dmc:
    """Device state of a DMC run over a Walkers set (qmc_dmc_state): the
    step weights, their block sums, the last branching's sources, the second
    generator (seeded from the 4 limbs setrn2 keeps) and the counters."""

    def __init__(self, walkers, limbs2):
        import torch
        self.wk = walkers
        nw, dev = walkers.nw, walkers.x.device
        self.wate = torch.ones(nw, dtype=torch.float64, device=dev)
        self.wsum = torch.zeros(nw, dtype=torch.float64, device=dev)
        self.src = torch.zeros(nw, dtype=torch.int32, device=dev)
        self.rng2 = torch.tensor([abi.limbs_to_state(limbs2)], dtype=torch.int64, device=dev)
        self.counts = torch.zeros(4, dtype=torch.int32, device=dev)
        s = abi.QmcDmcState()
        for name in ("wate", "wsum", "src", "rng2", "counts"):
            setattr(s, name, getattr(self, name).data_ptr())
        self.s = s

    def sweep(self, nsteps, delta, etrial, mmstep=1, mmult=10, mage=1e30, bm_order=0, stream=None):
        """nsteps DMC steps (qmc_dmc_sweep); the walkers' block sums get
        p_old times the weight, wsum the weights."""
        p = abi.QmcDmcParams()
        p.delta, p.etrial, p.mage = float(delta), float(etrial), float(mage)
        p.mmstep, p.mmult, p.bm_order, p.reserved0 = int(mmstep), int(mmult), int(bm_order), 0
        rc = lib().qmc_dmc_sweep(self.wk.ctx.h, C.byref(self.wk.s), C.byref(self.s), nsteps, C.byref(p),
                                 self.wk._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_dmc_sweep")


def dmc_branch(ctx, wate, rng2, src, counts, stream=None):
    """qmc_dmc_branch on device tensors (wate float64 [nw], overwritten;
    rng2 int64 [1]; src int32 [nw]; counts int32 [4])."""
    import torch
    st = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().qmc_dmc_branch(ctx.h, int(wate.numel()), C.c_void_p(wate.data_ptr()), C.c_void_p(rng2.data_ptr()),
                              C.c_void_p(src.data_ptr()), C.c_void_p(counts.data_ptr()),
                              C.c_void_p(st.cuda_stream))
    if rc:
        raise QmcError(rc, "qmc_dmc_branch")


def backflow_zero(deck):
    """True when every backflow table the deck's type pairs use is zero (the
    quasi-coordinates are the positions), as the library finds it."""
    nt = int(deck.ntypes)
    for it in range(nt):
        for jt in range(it, nt):
            k = int(deck.iub[it * 2 + jt])
            if k > 0 and np.any(deck.tables[k - 1].coef != 0.0):
                return False
    return True


class SingleMoves:
    """Single-electron moves over a Walkers set (qmc_sem_state): the inverse
    Slater matrices ([nw][2][mm][mm] complex, float64 or, with
    precision=abi.QMC_PREC_MIXED, float32; none for a deck with backflow,
    whose moves evaluate each proposal), the accumulated log|psi| change and
    the accepted-move count per walker."""

    def __init__(self, walkers, precision=0):
        import torch
        self.wk = walkers
        self.precision = int(precision)
        ctx = walkers.ctx
        np_ = [int(v) for v in ctx.deck.np[:ctx.deck.ntypes]]
        self.mm = max(np_)
        self.updates = backflow_zero(ctx.deck)
        nw, dev = walkers.nw, walkers.x.device
        rt = torch.float32 if self.precision == abi.QMC_PREC_MIXED else torch.float64
        shape = (nw, 2, self.mm, self.mm, 2) if self.updates else (1, 2)
        self.ainv = torch.zeros(shape, dtype=rt, device=dev)
        self.dlogpsi = torch.zeros(nw, dtype=torch.float64, device=dev)
        self.nacc = torch.zeros(nw, dtype=torch.float64, device=dev)
        s = abi.QmcSemState()
        s.ainv, s.dlogpsi, s.nacc = self.ainv.data_ptr(), self.dlogpsi.data_ptr(), self.nacc.data_ptr()
        self.s = s

    def refresh(self, stream=None):
        """The inverses from the walkers' current positions (qmc_sem_refresh)."""
        rc = lib().qmc_sem_refresh(self.wk.ctx.h, C.byref(self.wk.s), C.byref(self.s), self.precision,
                                   self.wk._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_sem_refresh")

    def moves(self, j0, j1, sigma, stream=None):
        """Moves of electrons j0..j1-1 in order (qmc_sem_moves)."""
        rc = lib().qmc_sem_moves(self.wk.ctx.h, C.byref(self.wk.s), C.byref(self.s), int(j0), int(j1),
                                 float(sigma), self.precision, self.wk._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_sem_moves")

    def sweep(self, sigma, stream=None):
        """One move of every electron, one spin per launch."""
        d = self.wk.ctx.deck
        bounds = [0] + [int(sum(d.np[:t + 1])) for t in range(d.ntypes)]
        for t in range(d.ntypes):
            if bounds[t + 1] > bounds[t]:
                self.moves(bounds[t], bounds[t + 1], sigma, stream)

    def advance(self, nmoves, sigma, stream=None):
        """nmoves moves per walker, continuing through the electrons in order
        from where the previous advance stopped (wrapping at N)."""
        n = int(self.wk.ctx.deck.nptot)
        cur = getattr(self, "_cursor", 0)
        left = int(nmoves)
        while left > 0:
            b = min(n, cur + left)
            self.moves(cur, b, sigma, stream)
            left -= b - cur
            cur = b % n
        self._cursor = cur

    def drift(self):
        """Largest relative distance, over walkers and spins, between the
        carried inverses and fresh FP64 inverses of the current positions
        (the carried ones are then replaced by the fresh ones: this is a
        refresh that also measures)."""
        import torch
        if not self.updates:
            return 0.0
        old = self.ainv.to(torch.float64, copy=True)
        self.refresh()
        new = self.ainv.to(torch.float64)
        num = (old - new).abs().amax(dim=(2, 3, 4))
        den = new.abs().amax(dim=(2, 3, 4)).clamp_min(1e-300)
        live = new.abs().amax(dim=(2, 3, 4)) > 0
        return float((num / den)[live].max().item())

    def set_tile(self, rows, threads_per_row):
        """qmc_sem_set_tile for this context and precision."""
        rc = lib().qmc_sem_set_tile(self.wk.ctx.h, self.precision, int(rows), int(threads_per_row))
        if rc:
            raise QmcError(rc, "qmc_sem_set_tile")

    def tune_tile(self, sigma, sweeps=3):
        """Autotune the move kernel's tile: for the row classes of this
        context's spins (32 or 64 rows), time `sweeps` sweeps with 2 and with
        4 threads per row on a copy of the walkers (the walkers themselves
        do not move) and keep the faster.  Returns {threads: ms per sweep}
        and the choice (None when every spin is in the 16-row class, which
        has one shape)."""
        import torch
        d = self.wk.ctx.deck
        rows = sorted({16 if m <= 16 else 32 if m <= 32 else 64 for m in d.np[:d.ntypes] if m > 0} - {16})
        if not rows or not self.updates:
            return {}, None
        wk = Walkers(self.wk.ctx, self.wk.x.cpu().numpy(), self.wk.rng.cpu().numpy())
        wk.start()
        sm = SingleMoves(wk, precision=self.precision)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        times = {}
        for cg in (2, 4):
            for r in rows:
                sm.set_tile(r, cg)
            sm.refresh()
            sm.sweep(sigma)
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(sweeps):
                sm.sweep(sigma)
            ev[1].record()
            torch.cuda.synchronize()
            times[cg] = ev[0].elapsed_time(ev[1]) / sweeps
        best = min(times, key=times.get)
        for r in rows:
            self.set_tile(r, best)
        return times, best

    def tune_refresh(self, sigma, candidates=(16, 64, 256, 1024), tol=1e-4, cycles=3, stream=None):
        """Refresh interval K (moves per walker between FP64 refreshes): for
        each candidate, `cycles` rounds of a refresh, K moves and a measured
        refresh (drift); returns the largest K whose worst drift stays
        within tol (the smallest candidate if none does) and the drifts.
        The default tol, 1e-4 relative on the largest entry of any walker's
        inverse, bounds the ratio R = (A^-1 a')_j to about that relative
        error; FP32 arithmetic alone leaves some 1e-6 there."""
        drifts = {}
        for k in sorted(candidates):
            self.refresh(stream)
            worst = 0.0
            for _ in range(cycles):
                self.advance(k, sigma, stream)
                worst = max(worst, self.drift())
            drifts[int(k)] = worst
        ok = [k for k in sorted(drifts) if drifts[k] <= tol]
        return (ok[-1] if ok else min(drifts)), drifts


def sem_moves_packed(moves, precision=0, stream=None):
    """qmc_sem_moves_packed: moves is a list of (SingleMoves, j0, j1, sigma),
    one per job (up to 16), all with the given precision."""
    import torch
    jobs = (abi.QmcSemJob * max(1, len(moves)))()
    for k, (sm, j0, j1, sigma) in enumerate(moves):
        if sm.precision != int(precision):
            raise ValueError("every packed job must use the call's precision")
        jobs[k].ctx = sm.wk.ctx.h.value
        jobs[k].w = C.cast(C.pointer(sm.wk.s), C.c_void_p).value
        jobs[k].d = C.cast(C.pointer(sm.s), C.c_void_p).value
        jobs[k].j0, jobs[k].j1, jobs[k].sigma = int(j0), int(j1), float(sigma)
    st = stream if stream is not None else torch.cuda.current_stream()
    rc = lib().qmc_sem_moves_packed(jobs, len(moves), int(precision), C.c_void_p(st.cuda_stream))
    if rc:
        raise QmcError(rc, "qmc_sem_moves_packed")


# fixed fields of one reptation trace row before the per-bead probabilities
RMC_TRACE_FIELDS = ("idir", "ndelta", "wnew", "age", "e", "ltf", "kin", "spare")


def rmc_params(delta, etrial, mage, mage_r, ntauskip, mdelta, driver=None, bm_order=0):
    """qmc_rmc_params from the rmc record and the deck's driver keywords
    (rejection, wpiu, gstorto, nodalaction, ecut; reptation.f:149-170)."""
    driver = driver or {}
    p = abi.QmcRmcParams()
    p.delta, p.etrial, p.mage, p.mage_r = delta, etrial, mage, mage_r
    p.ntauskip, p.mdelta = ntauskip, mdelta
    p.rejection = int(driver.get("rejection", ["1"])[0])
    p.wpiu = int(driver.get("wpiu", ["0"])[0])
    p.gstorto = int(driver.get("gstorto", ["0"])[0])
    p.nodalaction = int(driver.get("nodalaction", ["0"])[0])
    ec = driver.get("ecut", ["0"])
    p.ecut = int(ec[0])
    p.value_ecut = float(ec[1].replace("d", "e").replace("D", "E")) if p.ecut and len(ec) > 1 else 0.0
    p.bm_order = bm_order
    return p


class Reptiles:
    """Device reptile state of a batched reptation run (torch tensors on
    cuda).  Ring slots and path ends are 0-based."""

    def __init__(self, ctx, nr, ntau, mdelta, rng_states, block_sums=True):
        import torch
        self.ctx = ctx
        dev = torch.device("cuda", 0)
        N = ctx.deck.nptot
        NP = ctx.info["n_props"]
        f64 = torch.float64
        self.nr, self.ntau, self.mdelta = nr, ntau, mdelta
        nbuf = ntau + mdelta
        self.nbuf = nbuf
        self.x = torch.zeros((nr, nbuf, N, 2), dtype=f64, device=dev)
        self.g = torch.zeros((nr, nbuf, N, 2), dtype=f64, device=dev)
        self.props = torch.zeros((nr, nbuf, NP), dtype=f64, device=dev)
        self.age = torch.zeros((nr, nbuf), dtype=f64, device=dev)
        self.w = torch.zeros((nr, nbuf, 5), dtype=f64, device=dev)
        self.ends = torch.zeros((nr, 2), dtype=torch.int32, device=dev)
        self.rage = torch.zeros(nr, dtype=f64, device=dev)
        self.pp_old = torch.zeros((nr, NP), dtype=f64, device=dev)
        self.grad2 = torch.zeros(nr, dtype=f64, device=dev)
        self.rng = torch.tensor(np.asarray(rng_states, dtype=np.int64), device=dev)
        self.blk = torch.zeros((nr, NP), dtype=f64, device=dev) if block_sums else None
        s = abi.QmcReptiles()
        s.nr, s.ntau, s.nbuf = nr, ntau, nbuf
        for name in ("x", "g", "props", "age", "w", "ends", "rage", "pp_old", "rng", "blk"):
            t = getattr(self, name)
            setattr(s, name, t.data_ptr() if t is not None else None)
        s.grad2_ph = self.grad2.data_ptr()
        self.s = s

    def _stream(self, stream):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(st.cuda_stream)

    def start(self, beads, params, stream=None):
        """beads: [nr][ntau][N][2] initial paths (read_path)."""
        import torch
        self.x[:, :self.ntau] = torch.as_tensor(np.asarray(beads, dtype=np.float64),
                                                device=self.x.device)
        rc = lib().qmc_rmc_start(self.ctx.h, C.byref(self.s), C.byref(params), self._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_rmc_start")

    def sweep(self, nsteps, params, trace=False, stream=None):
        import torch
        tr = None
        if trace:
            tr = torch.zeros((nsteps, self.nr, len(RMC_TRACE_FIELDS) + self.mdelta),
                             dtype=torch.float64, device=self.x.device)
        rc = lib().qmc_rmc_sweep(self.ctx.h, C.byref(self.s), C.byref(params), nsteps,
                                 C.c_void_p(tr.data_ptr()) if trace else None, self._stream(stream))
        if rc:
            raise QmcError(rc, "qmc_rmc_sweep")
        return tr

    def status(self):
        """Sticky evaluation status per reptile (as Walkers.status)."""
        st = np.zeros(self.nr, np.int32)
        rc = lib().qmc_vmc_status(self.ctx.h, self.nr, abi.iptr(st))
        if rc:
            raise QmcError(rc, "qmc_vmc_status")
        return st
