"""ctypes binding of the CPU oracle (oracle/libqmc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package never
imports this module.
"""
import ctypes as C
import importlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libqmc_oracle.so")
# the same restatement evaluated in long double (64-bit significand): the
# exact value of the reference's formulas for the double inputs, against
# which the rounding error of an FP64 result is measured
LIB_LD = os.path.join(HERE, "libqmc_oracle_ld.so")

_pkg = importlib.import_module("2degas_amd")
abi = importlib.import_module("2degas_amd.abi")


def build(force=False):
    """Compile the C restatement with gcc (seconds), in double and in long
    double."""
    src = os.path.join(HERE, "qmc_oracle.c")
    for out, extra in ((LIB, []), (LIB_LD, ["-DORC_REAL=long double"])):
        if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
            subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared"] + extra +
                                  ["-o", out, src, "-lm"])
    return LIB


_libs = {}


def lib(extended=False):
    """The oracle library (extended=True: the long-double build)."""
    if extended not in _libs:
        path = LIB_LD if extended else LIB
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp, ip = abi.c_double_p, abi.c_int32_p
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.QmcTables), C.POINTER(C.c_int)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_info.argtypes = [C.c_void_p, C.POINTER(abi.QmcInfo)]
        L.orc_set_twist.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_twist_gvec.argtypes = [dp, C.c_int, dp, C.c_int, C.c_int, C.c_int, dp, C.c_int]
        L.orc_eval_batch.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, dp, ip]
        L.orc_vmc_replay.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, ip, C.c_int,
                                     dp, dp, dp, dp, dp, ip]
        L.orc_set_delta.argtypes = [C.c_void_p, C.c_double]
        L.orc_rannyu.restype = C.c_double
        _libs[extended] = L
    return _libs[extended]


def twist_gvec(kvec, el, ntheta, ith, jth, variant=0):
    kvec = np.ascontiguousarray(kvec, dtype=np.float64)
    nk = kvec.shape[0]
    el = np.ascontiguousarray(el, dtype=np.float64)
    out = np.zeros((2 * nk + 1, 2))
    lib().orc_twist_gvec(abi.dptr(kvec), nk, abi.dptr(el), ntheta, ith, jth,
                         abi.dptr(out), variant)
    return out


class Oracle:
    """compute_properties(1) on the CPU for a deck (one context).  With
    extended=True the sums run in long double: the exact value of the
    reference's formulas for the same double inputs (the twist list is
    still the double one, as the reference builds it)."""

    def __init__(self, deck, twist=(1, 1), gvec=None, extended=False):
        self.deck = deck
        self.L = lib(extended)
        s, self._keep = deck.pack()
        err = C.c_int(0)
        self.h = self.L.orc_create(C.byref(s), C.byref(err))
        if not self.h:
            raise ValueError(f"oracle rejected deck (code {err.value})")
        if gvec is None:
            gvec = twist_gvec(deck.kvec, deck.el, deck.ntheta, twist[0], twist[1], 0)
        self.gvec = np.ascontiguousarray(gvec, dtype=np.float64)
        self.L.orc_set_twist(self.h, abi.dptr(self.gvec), self.gvec.shape[0])
        inf = abi.QmcInfo()
        self.L.orc_info(self.h, C.byref(inf))
        self.info = inf.as_dict()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_delta(self, delta):
        self.L.orc_set_delta(self.h, float(delta))

    def eval(self, x, grad2_in=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, self.deck.nptot, 2)
        nw = x.shape[0]
        props = np.zeros((nw, self.info["n_props"]))
        g = np.zeros((nw, self.deck.nptot, 2))
        h = np.zeros((nw, 2))
        g2 = np.zeros(nw) if grad2_in is None else np.array(grad2_in, dtype=np.float64)
        st = np.zeros(nw, np.int32)
        rc = self.L.orc_eval_batch(self.h, nw, abi.dptr(x), abi.dptr(props), abi.dptr(g),
                                  abi.dptr(h), abi.dptr(g2), abi.iptr(st))
        if rc:
            raise RuntimeError(f"oracle eval failed: {rc}")
        return dict(props=props, g=g, h=h, grad2_ph=g2, status=st)

    def vmc_replay(self, x0, nsteps, delta, seed_limbs, bm_order=0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(self.deck.nptot, 2)
        p = np.zeros(nsteps); age = np.zeros(nsteps)
        ltf = np.zeros(nsteps); e = np.zeros(nsteps)
        xf = np.zeros_like(x0)
        sd = np.ascontiguousarray(seed_limbs, dtype=np.int32)
        sf = np.zeros(4, np.int32)
        rc = self.L.orc_vmc_replay(self.h, abi.dptr(x0), nsteps, delta, abi.iptr(sd), bm_order,
                                  abi.dptr(p), abi.dptr(age), abi.dptr(ltf), abi.dptr(e),
                                  abi.dptr(xf), abi.iptr(sf))
        if rc:
            raise RuntimeError(f"oracle replay failed: {rc}")
        return dict(p=p, age=age, ltf_new=ltf, e_new=e, x_final=xf, seed_final=sf)


def _pool_eval(args):
    deckdir, x, extended, twist, delta = args
    d = _pkg.deck.read_deck(deckdir)
    o = Oracle(d, twist=twist, extended=extended)
    if delta is not None:
        o.set_delta(delta)
    return o.eval(x)


def eval_parallel(deckdir, x, extended=False, nproc=None, twist=(1, 1), delta=None):
    """Oracle evaluation of many configurations in worker processes started
    fresh (spawn: no GPU state is inherited), nproc of them (default: the
    CPUs of this process, at most 16)."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    if nproc is None:
        nproc = max(1, min(16, len(os.sched_getaffinity(0))))
    x = np.ascontiguousarray(x, dtype=np.float64)
    chunks = np.array_split(x, min(len(x), nproc * 4))
    with ProcessPoolExecutor(nproc, mp_context=mp.get_context("spawn")) as ex:
        parts = list(ex.map(_pool_eval, [(deckdir, c, extended, tuple(twist), delta)
                                         for c in chunks]))
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def rank_seed_limbs(seed, rank):
    """RNG limbs of rank `rank` after input's seeding
    (reptation.f:37-39: seed(4)=2*(seed(4)+mytid)+1, setrn)."""
    s = list(seed)
    s[3] = 2 * (s[3] + rank) + 1
    l = [v % 4096 for v in s]
    l[3] = 2 * (l[3] // 2) + 1
    return np.array(l, np.int32)


# ---------------------------------------------------------------------------
# The reference itself (oracle/_ref, built by build_ref.sh from the reference
# sources in this container only).

REF_DIR = os.path.join(HERE, "_ref")


_LCG_A = 34522712143931
_MASK48 = (1 << 48) - 1


def branch2(wate, limbs2):
    """branch2 (reptation.f:2162-2271) on one rank holding the walkers in
    slots 1..n, restated: weights bounded at wmult = 10 (reptation.f:2176-2182),
    normalized to the population in slot order, integer parts kept and the
    fractional parts combed with nint(sum) draws of the second generator
    (reptation.f:2189-2210, the LCG x <- a x mod 2^48 of rannyu2,
    reptation.f:5356-5399), then every slot with multiplicity 0, in
    ascending order, takes a copy from the lowest slot with more than one
    (prov_dest, reptation.f:2258-2271).
    Returns (src, nmult, limbs2 after, mult): src[j] is the 0-based slot
    slot j's configuration comes from."""
    w = np.array(wate, dtype=np.float64)
    n = len(w)
    nmult = int(np.sum(w > 10.0))
    w = np.where(w > 10.0, 10.0, w)
    somma = 0.0
    for v in w:
        somma = somma + float(v)
    tm = np.zeros(n, dtype=np.int64)
    dl = np.zeros(n)
    k0 = 0.0
    for i in range(n):
        ww = (float(n) * float(w[i])) / somma
        tm[i] = int(ww)
        dl[i] = ww - float(tm[i])
        k0 = k0 + dl[i]
    x = abi.limbs_to_state(limbs2)
    j, k = 0, 1.0
    ndraw = int(np.floor(k0 + 0.5)) if k0 >= 0 else 0
    for _ in range(ndraw):
        x = (_LCG_A * x) & _MASK48
        r = x * 2.0 ** -48
        k = k - 1.0
        while k < r and j < n:
            j += 1
            k = k + dl[j - 1]
        tm[j - 1] += 1
        while k < 1.0 and j < n:
            j += 1
            k = k + dl[j - 1]
    mult = tm.copy()
    src = np.arange(n)
    m = tm.copy()
    for jj in range(n):
        if m[jj] == 0:
            for i in range(n):
                if m[i] >= 2:
                    src[jj] = i
                    m[i] -= 1
                    m[jj] += 1
                    break
    return src, nmult, abi.state_to_limbs(x), mult


def parse_branch(raw, n):
    """Mode 5 output: the slot origins (1-based), nmult and the second
    generator's limbs."""
    a = np.frombuffer(raw, dtype=np.int32)[5:]
    return {"src": a[:n] - 1, "nmult": int(a[n]), "limbs2": [int(v) for v in a[n + 1:n + 5]]}


def have_ref():
    return os.path.exists(os.path.join(REF_DIR, "refeval"))


def run_refeval(deckdir, workdir, mode, x, twist=(1, 1), delta=0.0, nsteps=None,
                seed=None, rmc=None, deck=None):
    """Run the reference harness on configurations x ([n][nptot][2]) in a
    copy of the deck; returns the raw output bytes.  Mode 4 (reptation)
    takes rmc = dict(ntau, ntauskip, mdelta, etrial, mage, mage_r, every)
    and the deck (for its per-type x_file names): x then holds the ntau
    initial beads, written as <x_file>.0 the way read_conf reads them."""
    import shutil
    os.makedirs(workdir, exist_ok=True)
    for fn in os.listdir(deckdir):
        shutil.copy(os.path.join(deckdir, fn), workdir)
    if seed is not None:
        with open(os.path.join(workdir, "seed"), "w") as f:
            f.write(" ".join(str(int(v)) for v in seed) + "\n")
    np.ascontiguousarray(x, dtype=np.float64).tofile(os.path.join(workdir, "cfg.bin"))
    n = nsteps if nsteps is not None else x.shape[0]
    with open(os.path.join(workdir, "refeval.ctl"), "w") as f:
        f.write(f"{mode}\n{twist[0]} {twist[1]}\n{n}\n{delta!r}\ncfg.bin\nout.bin\n")
        if rmc is not None:
            f.write("{ntau} {ntauskip} {mdelta} {etrial!r} {mage!r} {mage_r!r} {every}\n"
                    .format(**{k: (float(v) if k in ("etrial", "mage", "mage_r") else int(v))
                               for k, v in rmc.items()}))
    if rmc is not None:
        p0 = 0
        for t in range(deck.ntypes):
            npt = deck.np[t]
            with open(os.path.join(workdir, deck.x_file[t] + ".0"), "w") as f:
                for b in range(int(rmc["ntau"])):
                    for ip in range(p0, p0 + npt):
                        f.write(f"{float(x[b, ip, 0])!r} {float(x[b, ip, 1])!r}\n")
            p0 += npt
    r = subprocess.run([os.path.join(REF_DIR, "refeval")], cwd=workdir,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(r.stdout.decode()[-2000:])
    with open(os.path.join(workdir, "out.bin"), "rb") as f:
        return f.read(), r.stdout.decode()


def read_gvec_dump(workdir):
    """g-vectors the harness used (twistg or input's list)."""
    with open(os.path.join(workdir, "gvec.bin"), "rb") as f:
        b = f.read()
    n = int(np.frombuffer(b[:4], dtype=np.int32)[0])
    return np.frombuffer(b[4:4 + 16 * n], dtype=np.float64).reshape(n, 2).copy()


def parse_eval(raw, nptot):
    head = np.frombuffer(raw[:20], dtype=np.int32)
    ns, npt, nt, nk, nr = [int(v) for v in head]
    assert npt == nptot
    per = ns + 2 * npt + nt + 1 + nr * 2 * nk
    d = np.frombuffer(raw[20:], dtype=np.float64).reshape(-1, per)
    o = 0
    out = {"scalars": d[:, o:o + ns]}; o += ns
    out["g"] = d[:, o:o + 2 * npt].reshape(-1, npt, 2); o += 2 * npt
    out["h"] = d[:, o:o + nt]; o += nt
    out["grad2_ph"] = d[:, o]; o += 1
    out["rhok"] = d[:, o:o + nr * 2 * nk]
    return out


def parse_rmc(raw, nptot, nsteps, every, mdelta):
    """Mode 4 stream: per checkpoint the full reptile ring (1-based bead
    indices as the reference keeps them), per move idir, ndelta, path weight,
    bead acceptance probabilities, and after the test age and the path
    averages of E, -log psi and the kinetic energy."""
    o = 20
    nps, nbuf, nprops = [int(v) for v in np.frombuffer(raw[o:o + 12], dtype=np.int32)]
    o += 12

    def take(n, dt=np.float64):
        nonlocal o
        sz = np.dtype(dt).itemsize * n
        a = np.frombuffer(raw[o:o + sz], dtype=dt).copy()
        o += sz
        return a

    ck = {k: [] for k in ("seed", "jfirst", "jlast", "age", "grad2_ph", "x", "g", "p",
                          "age_stack", "w", "s", "pp_old")}
    mv = {k: [] for k in ("idir", "ndelta", "wnew", "pbead", "age", "e", "ltf", "kin")}
    for s in range(nsteps):
        if s % every == 0:
            h = take(6, np.int32)
            ck["seed"].append(h[:4]); ck["jfirst"].append(h[4]); ck["jlast"].append(h[5])
            a = take(2)
            ck["age"].append(a[0]); ck["grad2_ph"].append(a[1])
            ck["x"].append(take(nbuf * nptot * 2).reshape(nbuf, nptot, 2))
            ck["g"].append(take(nbuf * nptot * 2).reshape(nbuf, nptot, 2))
            ck["p"].append(take(nbuf * nps).reshape(nbuf, nps))
            ck["age_stack"].append(take(nbuf))
            ck["w"].append(take(5 * nbuf).reshape(5, nbuf).T.copy())
            ck["s"].append(take(nbuf))
            ck["pp_old"].append(take(nprops))
        h = take(2, np.int32)
        mv["idir"].append(h[0]); mv["ndelta"].append(h[1])
        r = take(1 + mdelta)
        mv["wnew"].append(r[0]); mv["pbead"].append(r[1:])
        r = take(4)
        mv["age"].append(r[0]); mv["e"].append(r[1]); mv["ltf"].append(r[2]); mv["kin"].append(r[3])
    fin = take(6, np.int32)
    out = {"ck_" + k: np.array(v) for k, v in ck.items()}
    out.update({k: np.array(v) for k, v in mv.items()})
    out.update(nps=nps, nbuf=nbuf, nprops=nprops, seed_final=fin[:4], jfirst_final=fin[4],
               jlast_final=fin[5])
    return out


def parse_replay(raw, nptot, nsteps):
    o = 20
    seeds = np.zeros((nsteps, 4), np.int32)
    xold = np.zeros((nsteps, nptot, 2))
    rec = np.zeros((nsteps, 6))
    for s in range(nsteps):
        seeds[s] = np.frombuffer(raw[o:o + 16], dtype=np.int32); o += 16
        xold[s] = np.frombuffer(raw[o:o + 16 * nptot], dtype=np.float64).reshape(nptot, 2)
        o += 16 * nptot
        rec[s] = np.frombuffer(raw[o:o + 48], dtype=np.float64); o += 48
    xf = np.frombuffer(raw[o:o + 16 * nptot], dtype=np.float64).reshape(nptot, 2); o += 16 * nptot
    sf = np.frombuffer(raw[o:o + 16], dtype=np.int32)
    return dict(seeds=seeds, x_old=xold, p=rec[:, 0], age=rec[:, 1], ltf_new=rec[:, 2], e_new=rec[:, 3],
                ltf_old=rec[:, 4], e_old=rec[:, 5], x_final=xf, seed_final=sf)
