"""Batched VMC and RQMC runs with the reference's inputs and outputs.

`run(dirpath, nw)` does for the `vmc` and `rmc` commands of `<runid>.in`
what the reference's driver does (sonaseppia, reptation.f:712-807; vmc,
reptation.f:973-1010; rmc, reptation.f:1286-1313), for a batch of nw
walkers (or reptiles) on the GPU, walker k playing MPI rank k of the
reference:

* inputs: `runid`; `seed` (4 limbs; rank k draws from the limbs
  (s1, s2, s3, 2 (s4 + k) + 1), the second generator keeps (s1..s4),
  reptation.f:25-39); the deck; the configurations `<x_file>.<k>`, or
  `<x_file>.res.<k>` after a `restart` record (read_conf,
  reptation.f:2452-2495);
* after every block (averages(3), reptation.f:4306-4434): the block and
  cumulative averages of the scalar properties with the cumulative error on
  stdout, `<runid>.rhok_<type>` for each rhok type (write_nonscalar_props,
  reptation.f:4586-4615), the restart file `<runid>.res` and
  `<x_file>.res.<k>` (restart(1), reptation.f:4436-4509);
* a `restart` record reads `<runid>.res` back (restart(0),
  reptation.f:4510-4536): the next block, the cumulative sums, every rank's
  generator state and, with twists, the current twist;
* at the end rank 0's generator state goes to `seed` (reptation.f:844-851);
* at the end of the command every rank's configuration goes back to
  `<x_file>.<k>` (write_conf, reptation.f:1008-1011, 2503-2535);
* a deck with `ntwists n` runs the blocks once per twist (ith, jth) in
  1..n x 1..n, as tabc does (reptation.f:9559-9638), each twist reading
  the configurations the previous one wrote;
* an `rmc` record reads each rank's initial path of ntau configurations
  from `<x_file>.<k>` (read_path, reptation.f:1766-1802), prints
  `rmc: nprops =`, resets the reptile age at block 1 (averages(1),
  reptation.f:4290), writes `===>> rmc block` lines, `.res` with the path
  from jfirst in `<x_file>.res.<k>` after every block (restart(1),
  reptation.f:4481-4490) and the final path to `<x_file>.<k>` (write_path,
  reptation.f:1741-1764); after a `restart` record the trial energy is the
  cumulative mean energy of the restart file (restart(0), reptation.f:4525);
* a `dmc` record (reptation.f:768-782; dmc, reptation.f:872-921) runs
  diffusion Monte Carlo with branch2's fixed population, each rank holding
  one configuration (ntarget 1): per step mmstep moves with their weights,
  the weighted block sums, the branching of every rank's walker; after each
  block the `===>> dmc block` line with nconf, the population bounds and
  rank 0's forced moves, the trial energy set to the cumulative mean energy,
  `.res` (both generators of every rank) and `<x_file>.res.<k>`; at the end
  `<x_file>.<k>` and the count of weights bounded on rank 0.

Block averages sum p_old over the walkers and steps of the block, the
MPI_REDUCE of the reference (reptation.f:4308-4311), so that blk_norm is
nw * nstp.  Numbers go out with the reference's edit descriptors where it
uses them (2e19.11, e9.2, e10.3); list-directed items are written as
shortest round-trip reals, which Fortran list-directed input reads back.
The evaluation, the moves and the block sums run on the device through the
C-ABI (qmc_vmc_start / qmc_vmc_sweep); this module is host bookkeeping.
"""
import math
import os
import sys

import numpy as np

from . import abi, deck as deckmod, hip


def fortran_e(v, w, d):
    """The Fortran edit descriptor Ew.d: 0.<d digits>E+xx right-justified in
    w columns (three-digit exponents drop the E, as Fortran does)."""
    v = float(v)
    if v == 0.0 or not math.isfinite(v):
        body = "0." + "0" * d + "E+00" if v == 0.0 else ("NaN" if v != v else ("-Inf" if v < 0 else "Inf"))
        return body.rjust(w)
    s = f"{abs(v):.{d - 1}e}"          # m.mmmme+xx, d significant digits
    mant, ex = s.split("e")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    sign = "-" if v < 0 else ""
    if abs(e) <= 99:
        tail = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        tail = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return f"{sign}0.{digits}{tail}".rjust(w)


def _ld(v):
    """A real for list-directed output: shortest round-trip form."""
    return repr(float(v))


def read_seed(dirpath):
    """The 4 seed limbs, (0, 0, 0, 1) when the file is missing or empty
    (reptation.f:25, 31-33)."""
    p = os.path.join(dirpath, "seed")
    try:
        with open(p) as f:
            v = [int(t) for t in f.read().split()[:4]]
    except FileNotFoundError:
        v = []
    return v if len(v) == 4 else [0, 0, 0, 1]


def setrn_limbs(seed):
    """The limbs setrn keeps: each mod 4096, the last made odd
    (reptation.f:5292-5299)."""
    l = [int(v) % 4096 for v in seed]
    l[3] = 2 * (l[3] // 2) + 1
    return l


def rank_limbs(seed, k):
    """Generator limbs of rank k (reptation.f:37-38)."""
    return setrn_limbs([seed[0], seed[1], seed[2], 2 * (seed[3] + k) + 1])


def read_script(dirpath, runid):
    """Records of `<runid>.in` as word lists, up to `end`."""
    recs = []
    with open(os.path.join(dirpath, runid + ".in")) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            recs.append(w)
            if w[0] == "end":
                break
    return recs


def scalar_names(deck, info):
    """Names of the scalar property slots (input, reptation.f:577-616)."""
    names = [""] * info["n_scalar_props"]
    fixed = {"jetot": "elocal", "jltf": "-log tf", "jacc": "acc.rate", "jpot0": "epot",
             "jkin0": "ekin", "je2": "e2", "jgg": "gg", "jefp": "efp"}
    for key, nm in fixed.items():
        names[info[key] - 1] = nm
    for it, tname in enumerate(deck.typenames):
        for key, pre in (("jpot", "pot_"), ("jkin", "kin_"), ("jun", "un_")):
            j = info[key][it]
            if j > 0:
                names[j - 1] = pre + tname
    return names


class _Stats:
    """Cumulative sums of averages (cml_av, cml2, cml_norm)."""

    def __init__(self, n_props):
        self.av = np.zeros(n_props)
        self.sq = np.zeros(n_props)
        self.norm = 0.0

    def add(self, blk, blk_norm):
        self.av += blk
        self.sq += blk * blk / blk_norm
        self.norm += blk_norm

    def err(self, iblk):
        if iblk <= 1:
            return np.zeros_like(self.av)
        return np.sqrt(np.abs((self.sq / self.norm - (self.av / self.norm) ** 2) / (iblk - 1)))


class Run:
    """State of a batched VMC run in the reference's directory layout."""

    def __init__(self, dirpath, nw, out=None):
        self.dir = dirpath
        self.nw = nw
        self.out = out or sys.stdout
        with open(os.path.join(dirpath, "runid")) as f:
            self.runid = f.read().split()[0]
        self.deck = deckmod.read_deck(dirpath, self.runid)
        self.seed = read_seed(dirpath)
        # generator limbs per rank: primary (per rank) and second (shared)
        self.limbs = [rank_limbs(self.seed, k) for k in range(nw)]
        self.limbs2 = [setrn_limbs(self.seed) for _ in range(nw)]
        self.res_string = "."
        self.iblk0 = 1
        self.stats = None
        self.twist = (1, 1)

    def _path(self, name):
        return os.path.join(self.dir, name)

    # -- files
    def read_configs(self, nconf=1):
        """nconf configurations per rank from `<x_file><res_string><k>`
        (read_conf, reptation.f:2452-2495): [nw][N][2], or [nw][nconf][N][2]
        when nconf > 1."""
        d = self.deck
        x = []
        for k in range(self.nw):
            paths = [self._path(f + self.res_string + str(k)) for f in d.x_file]
            c = deckmod.read_configs(paths, d.np, nconf)
            if len(c) < nconf:
                raise ValueError(f"rank {k}: {len(c)} configurations in {paths}, {nconf} needed")
            x.append(c[0] if nconf == 1 else c[:nconf])
        return np.stack(x)

    def write_paths(self, x, ends, mid):
        """Each rank's path from jfirst, ntau beads, in `<x_file><mid><k>`
        (write_path, reptation.f:1741-1764; restart(1) for rmc,
        reptation.f:4481-4490): per type file, the beads in path order."""
        d = self.deck
        nr, nbuf = x.shape[0], x.shape[1]
        off = 0
        for it, f in enumerate(d.x_file):
            n = d.np[it]
            for k in range(nr):
                with open(self._path(f"{f}{mid}{k}"), "w") as fh:
                    for i in range(self.ntau):
                        b = (int(ends[k, 0]) + i) % nbuf
                        for p in x[k, b, off:off + n]:
                            fh.write(f" {_ld(p[0])} {_ld(p[1])}\n")
            off += n

    def write_configs(self, x, mid=".res."):
        """`<x_file><mid><k>`: `.res.` after every block (restart(1)), `.`
        at the end of a vmc command (write_conf, reptation.f:2503-2535)."""
        d = self.deck
        off = 0
        for it, f in enumerate(d.x_file):
            n = d.np[it]
            for k in range(self.nw):
                with open(self._path(f"{f}{mid}{k}"), "w") as fh:
                    for p in x[k, off:off + n]:
                        fh.write(f" {_ld(p[0])} {_ld(p[1])}\n")
            off += n

    def write_restart(self, iblk, who="vmc"):
        st = self.stats
        with open(self._path(self.runid + ".res"), "w") as f:
            f.write(f" {iblk}  {who}\n")
            f.write(f" {_ld(st.norm)}\n")
            for a, b in zip(st.av, st.sq):
                f.write(f" {_ld(a)} {_ld(b)}\n")
            for k in range(self.nw):
                f.write(" " + " ".join(str(v) for v in self.limbs[k] + self.limbs2[k]) + "\n")
            if self.deck.ntheta != 0:
                f.write(f" {self.twist[0]} {self.twist[1]}\n")

    def read_restart(self, nblk, n_props, who="vmc"):
        with open(self._path(self.runid + ".res")) as f:
            tok = f.read().split()
        iblk = int(tok[0])
        pos = 2
        self.iblk0 = iblk % nblk + 1
        st = _Stats(n_props)
        st.norm = deckmod._ffloat(tok[pos])
        pos += 1
        for i in range(n_props):
            st.av[i] = deckmod._ffloat(tok[pos])
            st.sq[i] = deckmod._ffloat(tok[pos + 1])
            pos += 2
        nrank = min(self.nw, (len(tok) - pos) // 8)
        if nrank < self.nw:
            raise ValueError(f"{self.runid}.res holds {nrank} ranks, the run has {self.nw}")
        for k in range(self.nw):
            v = [int(t) for t in tok[pos:pos + 8]]
            pos += 8
            self.limbs[k], self.limbs2[k] = v[:4], v[4:]
        if self.deck.ntheta != 0 and pos + 1 < len(tok):
            self.twist = (int(tok[pos]), int(tok[pos + 1]))
        self.stats = st

    def write_seed(self):
        with open(self._path("seed"), "w") as f:
            f.write(" " + " ".join(str(v) for v in self.limbs[0]) + "\n")

    # -- output of a block
    def block_output(self, iblk, blk, blk_norm, names, info, who="vmc", extra=""):
        st = self.stats
        err = st.err(iblk)
        w = self.out.write
        w(f" ===>> {who} block  {iblk}{extra}\n")
        for i, nm in enumerate(names):
            w(fortran_e(blk[i] / blk_norm, 19, 11) + fortran_e(st.av[i] / st.norm, 19, 11)
              + fortran_e(err[i], 9, 2) + fortran_e(blk_norm, 10, 3) + "  " + nm.ljust(20) + "\n")
        d = self.deck
        nk2 = 2 * len(np.asarray(d.kvec).reshape(-1, 2)) if d.nrhok else 0
        for j in range(1, d.nrhok + 1):
            t = d.irhok.index(j)
            i0 = info["jrhok"] - 1 + (j - 1) * nk2
            with open(self._path(f"{self.runid}.rhok_{d.typenames[t]}"), "w") as f:
                for n in range(nk2):
                    i = i0 + n
                    f.write(fortran_e(blk[i] / blk_norm, 19, 11) + fortran_e(st.av[i] / st.norm, 19, 11)
                            + fortran_e(err[i], 9, 2) + fortran_e(blk_norm, 10, 3) + " " + " merea "
                            + f"{n + 1:5d}" + "\n")

    # -- the vmc command
    def vmc(self, nblk, nstp, delta):
        import torch
        d = self.deck
        twists = [(1, 1)]
        if d.ntheta != 0:
            if self.res_string != ".":
                i0, j0 = self.twist if self.iblk0 != 1 else \
                    (self.twist[0] + self.twist[1] // d.ntheta, self.twist[1] % d.ntheta + 1)
            else:
                i0, j0 = 1, 1
            twists = []
            for ith in range(i0, d.ntheta + 1):
                for jth in range(j0 if ith == i0 else 1, d.ntheta + 1):
                    twists.append((ith, jth))
        mage = float(nstp * nblk)
        for tw in twists:
            self.twist = tw
            ctx = hip.Context(d, twist=tw)
            try:
                info = ctx.info
                names = scalar_names(d, info)
                x = self.read_configs()
                states = [abi.limbs_to_state(l) for l in self.limbs]
                wk = hip.Walkers(ctx, x, states)
                wk.start(delta=delta)
                for iblk in range(self.iblk0, nblk + 1):
                    if iblk == 1 or self.stats is None:
                        self.stats = _Stats(info["n_props"])
                    wk.blk.zero_()
                    wk.sweep(nstp, delta, mage=mage)
                    torch.cuda.synchronize()
                    st = wk.status()
                    if np.any(st & 2):
                        raise RuntimeError("singular orbital matrix (the reference stops here, reptation.f:10318)")
                    blk = wk.blk.sum(dim=0).cpu().numpy()
                    blk_norm = float(self.nw * nstp)
                    self.stats.add(blk, blk_norm)
                    self.limbs = [abi.state_to_limbs(int(v)) for v in wk.rng.cpu().numpy()]
                    self.block_output(iblk, blk, blk_norm, names, info)
                    self.write_restart(iblk)
                    self.write_configs(wk.x.cpu().numpy())
                # the end of vmc: the configurations go back to the files the
                # next command (or twist) reads, and after a restart the
                # following twists start at block 1 (reptation.f:1008-1015)
                self.write_configs(wk.x.cpu().numpy(), ".")
                if d.ntheta != 0 and self.res_string != ".":
                    self.iblk0 = 1
                wk = None
            finally:
                ctx.close()

    # -- the rmc command
    def rmc(self, nblk, nstp, delta, ntau, ntauskip, mdelta, etrial, mage, mage_r):
        import torch
        d = self.deck
        self.ntau = ntau
        # rmc runs on the g-vectors input set up, without the twist loop
        # (sonaseppia calls rmc directly, reptation.f:795-807)
        gv = hip.twist_gvec(d.kvec, d.el, 0, 0, 0)
        ctx = hip.Context(d, gvec=gv)
        try:
            info = ctx.info
            names = scalar_names(d, info)
            if self.res_string != ".":
                self.read_restart(nblk, info["n_props"], who="rmc")
                # restart(0) takes the trial energy from the cumulative
                # average (reptation.f:4525)
                etrial = self.stats.av[info["jetot"] - 1] / self.stats.norm
            prm = hip.rmc_params(delta, etrial, mage, mage_r, ntauskip, mdelta, d.driver)
            beads = self.read_configs(ntau)
            states = [abi.limbs_to_state(l) for l in self.limbs]
            rep = hip.Reptiles(ctx, self.nw, ntau, mdelta, states)
            rep.start(beads, prm)
            self.out.write(f" rmc: nprops = {info['n_props']:12d}\n")
            for iblk in range(self.iblk0, nblk + 1):
                if iblk == 1 or self.stats is None:
                    self.stats = _Stats(info["n_props"])
                    # averages(1) at block 1 resets the reptile age
                    # (reptation.f:4286-4290)
                    rep.rage.zero_()
                rep.blk.zero_()
                rep.sweep(nstp, prm)
                torch.cuda.synchronize()
                st = rep.status()
                if np.any(st & 2):
                    raise RuntimeError("singular orbital matrix (the reference stops here, reptation.f:10318)")
                blk = rep.blk.sum(dim=0).cpu().numpy()
                blk_norm = float(self.nw * nstp)
                self.stats.add(blk, blk_norm)
                self.limbs = [abi.state_to_limbs(int(v)) for v in rep.rng.cpu().numpy()]
                self.block_output(iblk, blk, blk_norm, names, info, who="rmc")
                self.write_restart(iblk, who="rmc")
                self.write_paths(rep.x.cpu().numpy(), rep.ends.cpu().numpy(), ".res.")
            self.write_paths(rep.x.cpu().numpy(), rep.ends.cpu().numpy(), ".")
            rep = None
        finally:
            ctx.close()

    # -- the dmc command
    def dmc(self, nblk, nstp, delta, ntarget, mmstep, etrial, mage, mmult):
        import torch
        d = self.deck
        if ntarget != 1:
            raise NotImplementedError("dmc: each rank holds one configuration here (ntarget 1); the "
                                      "reference's ranks with several share one random stream, "
                                      "walker after walker")
        if d.ntheta != 0:
            raise NotImplementedError("dmc with twist averaging (tabc(dmc)) is not on this path")
        ctx = hip.Context(d, twist=(1, 1))
        try:
            info = ctx.info
            names = scalar_names(d, info)
            jetot = info["jetot"] - 1
            if self.res_string != ".":
                self.read_restart(nblk, info["n_props"], who="dmc")
                etrial = self.stats.av[jetot] / self.stats.norm
            x = self.read_configs()
            states = [abi.limbs_to_state(l) for l in self.limbs]
            wk = hip.Walkers(ctx, x, states)
            wk.start(delta=delta)
            dm = hip.Dmc(wk, self.limbs2[0])
            for iblk in range(self.iblk0, nblk + 1):
                if iblk == 1 or self.stats is None:
                    self.stats = _Stats(info["n_props"])
                wk.blk.zero_()
                dm.wsum.zero_()
                # averages(1) resets rank 0's count of forced moves
                dm.counts[3] = 0
                dm.sweep(nstp, delta, etrial, mmstep=mmstep, mmult=mmult, mage=mage)
                torch.cuda.synchronize()
                st = wk.status()
                if np.any(st & 2):
                    raise RuntimeError("singular orbital matrix (the reference stops here, reptation.f:10318)")
                blk = wk.blk.sum(dim=0).cpu().numpy()
                blk_norm = float(dm.wsum.sum().item())
                cnt = dm.counts.cpu().numpy()
                self.stats.add(blk, blk_norm)
                self.limbs = [abi.state_to_limbs(int(v)) for v in wk.rng.cpu().numpy()]
                l2 = abi.state_to_limbs(int(dm.rng2.cpu().numpy()[0]))
                self.limbs2 = [l2 for _ in range(self.nw)]
                self.out.write(f" ***** mytid 0 nsg =  {_ld(0.0)}\n")
                # averages(3) for dmc: the trial energy is the cumulative
                # mean energy (reptation.f:4327-4333); min/max_nconf change
                # only in branch (dmc1), so they keep averages(1)'s values
                etrial = self.stats.av[jetot] / self.stats.norm
                self.block_output(iblk, blk, blk_norm, names, info, who="dmc",
                                  extra=f" nconf =  1  mnnc, mxnc =  10000000  0  nage =  {_ld(float(cnt[3]))}")
                self.write_restart(iblk, who="dmc")
                self.write_configs(wk.x.cpu().numpy())
            self.write_configs(wk.x.cpu().numpy(), ".")
            cnt = dm.counts.cpu().numpy()
            self.out.write(f" max.multiplicity  {mmult}  exceeded  {int(cnt[2])}  times\n")
            wk = dm = None
        finally:
            ctx.close()

    def execute(self):
        """Run the records of `<runid>.in`."""
        recs = read_script(self.dir, self.runid)
        n = 0
        while n < len(recs):
            w = recs[n]
            n += 1
            if w[0] == "end":
                break
            if w[0] == "restart":
                # read from .res files; the next j - 1 records are skipped
                # (reptation.f:730-735)
                self.res_string = ".res."
                n += int(w[1]) - 1 if len(w) > 1 else 0
                continue
            if w[0] == "vmc":
                nblk, nstp = int(w[1]), int(w[2])
                delta = deckmod._ffloat(w[3])
                if self.res_string != ".":
                    self.read_restart(nblk, self._n_props())
                self.vmc(nblk, nstp, delta)
                continue
            if w[0] == "rmc":
                # rmc nblk nstp delta ntau ntauskip mdelta etrial mage mage_r
                # (reptation.f:795-805)
                nblk, nstp = int(w[1]), int(w[2])
                delta = deckmod._ffloat(w[3])
                ntau, ntauskip, mdelta = int(w[4]), int(w[5]), int(w[6])
                etrial, mage, mage_r = (deckmod._ffloat(v) for v in w[7:10])
                self.rmc(nblk, nstp, delta, ntau, ntauskip, mdelta, etrial, mage, mage_r)
                continue
            if w[0] == "dmc":
                # dmc nblk nstp delta ntarget mmstep etrial mage mmult
                # (reptation.f:768-777)
                nblk, nstp = int(w[1]), int(w[2])
                delta = deckmod._ffloat(w[3])
                ntarget, mmstep = int(w[4]), int(w[5])
                etrial, mage = deckmod._ffloat(w[6]), deckmod._ffloat(w[7])
                mmult = int(w[8])
                self.dmc(nblk, nstp, delta, ntarget, mmstep, etrial, mage, mmult)
                continue
            raise NotImplementedError(f"{w[0]}: only vmc, rmc, dmc, restart and end records run on this path")
        self.write_seed()

    def _n_props(self):
        ctx = hip.Context(self.deck, twist=self.twist)
        try:
            return ctx.info["n_props"]
        finally:
            ctx.close()


def run(dirpath, nw, out=None):
    """Run `<runid>.in` in dirpath with nw walkers (ranks); returns the Run."""
    r = Run(dirpath, nw, out)
    r.execute()
    return r


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="batched VMC with the reference's files")
    ap.add_argument("dir")
    ap.add_argument("--nw", type=int, required=True, help="walkers (ranks)")
    a = ap.parse_args()
    run(a.dir, a.nw)
