"""Reader for the reference's input decks.

Reads what `input` (reptation.f:12-710) reads for the walker-evaluation
path: the `runid` and `seed` files, the `<runid>.sy` keyword deck
(readwords, reptation.f:4790-4842), spline tables with an optional k-space
block (t_read, reptation.f:4689-4738) and the k-vector list (kread,
reptation.f:4740-4769).  The result is packed into the `qmc_tables` struct
of the C-ABI.
"""
import os
from dataclasses import dataclass, field

import numpy as np

from .abi import QmcTables, dptr, iptr

# keywords of the .sy deck that only configure drivers outside this path;
# they are accepted and recorded (reptation.f:149-178)
_DRIVER_KEYWORDS = {"adrift", "gstorto", "wpiu", "nodalaction", "rejection",
                    "fullprop", "ecut"}
# keywords the evaluation path does not implement (dead or out-of-scope code)
_UNSUPPORTED = {"gofr", "mstar", "cmass", "itc", "sites", "vext", "vps",
                "bose-nosanow", "u3_spp", "backflow2", "mathieu", "plane-wave",
                "slater-nosanow", "nosanow", "lcao", "inc", "der"}


def _ffloat(s):
    return float(s.replace("D", "E").replace("d", "e"))


def read_words(lines):
    """Split deck records into word lists, honouring the reference's 80-char
    records and backslash continuation (readwords, reptation.f:4790-4842)."""
    recs = []
    cur = []
    for raw in lines:
        rec = raw.rstrip("\n")[:80]
        cont = False
        if "\\" in rec:
            rec = rec[: rec.index("\\")]
            cont = True
        cur.extend(w[:48] for w in rec.split())
        if not cont:
            recs.append(cur)
            cur = []
    if cur:
        recs.append(cur)
    return recs


@dataclass
class Table:
    name: str
    ngrid: int
    drt: float
    coef: np.ndarray           # [ngrid+1, 4]
    routine: str = ""
    params: list = field(default_factory=list)
    vk: np.ndarray = None      # [nk_ewald] or None
    vk0: float = 0.0
    tail: float = 0.0
    iexp: int = 0


def read_table(path):
    """t_read (reptation.f:4689-4738)."""
    with open(path) as f:
        toks = f.read().split("\n")
    it = iter(toks)
    head = next(it).split()
    ngrid, drt = int(head[0]), _ffloat(head[1])
    vals = []
    while len(vals) < 4 * (ngrid + 1):
        vals.extend(_ffloat(v) for v in next(it).split())
    coef = np.array(vals[: 4 * (ngrid + 1)], dtype=np.float64).reshape(ngrid + 1, 4)
    t = Table(os.path.basename(path), ngrid, drt, coef)
    rest = [ln for ln in it]
    k = 0
    while k < len(rest) and not rest[k].split():
        k += 1
    if k < len(rest):
        t.routine = rest[k].split()[0]
        k += 1
        npar = int(rest[k].split()[0])
        k += 1
        t.params = [rest[k + j].split() for j in range(npar)]
        k += npar
        while k < len(rest) and not rest[k].split():
            k += 1
        if k < len(rest) and rest[k].split()[0] == "kspace":
            nke = int(rest[k + 1].split()[0])
            vv = []
            j = k + 2
            while len(vv) < nke + 1:
                vv.extend(_ffloat(v) for v in rest[j].split())
                j += 1
            t.vk = np.array(vv[:nke], dtype=np.float64)
            t.vk0 = vv[nke]
    return t


def read_kvec(path, maxk=1 << 20):
    """kread (reptation.f:4740-4769): one k-vector per line."""
    ks = []
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if len(w) < 2:
                break
            ks.append((_ffloat(w[0]), _ffloat(w[1])))
            if len(ks) >= maxk:
                break
    return np.array(ks, dtype=np.float64).reshape(-1, 2)


def _zeros4():
    return np.zeros(4, np.int32)


def _nok():
    return np.zeros((0, 2))


@dataclass
class Deck:
    """State of `input` relevant to compute_properties."""
    dirpath: str
    runid: str
    seed: tuple
    ntypes: int = 0
    typenames: list = field(default_factory=list)
    np: list = field(default_factory=list)
    hbs2m: list = field(default_factory=list)
    x_file: list = field(default_factory=list)
    el: tuple = (0.0, 0.0)
    tables: list = field(default_factory=list)
    drt: float = 0.0
    iu2: object = field(default_factory=_zeros4)
    iu3: object = field(default_factory=_zeros4)
    iub: object = field(default_factory=_zeros4)
    iv2: object = field(default_factory=_zeros4)
    v0: float = 0.0
    kvec: object = field(default_factory=_nok)
    irhok: list = field(default_factory=lambda: [0, 0])
    nrhok: int = 0
    ntheta: int = 0
    adrift: float = 0.0
    driver: dict = field(default_factory=dict)

    @property
    def nptot(self):
        return int(sum(self.np))

    def _table(self, fname):
        for i, t in enumerate(self.tables):
            if t.name == os.path.basename(fname):
                return i + 1
        t = read_table(os.path.join(self.dirpath, fname))
        self.tables.append(t)
        self.drt = t.drt  # one global drt, last read wins (reptation.f:4705)
        return len(self.tables)

    def _type(self, name):
        return self.typenames.index(name)

    def pack(self):
        """Build the qmc_tables struct; returns (struct, keepalive list)."""
        ntab = len(self.tables)
        rows = max((t.ngrid + 1 + (len(t.vk) + 1 if t.vk is not None else 0)
                    for t in self.tables), default=1)
        ldut = rows + 1
        ut = np.zeros((max(ntab, 1), 4, ldut), dtype=np.float64)
        ngrid = np.zeros(max(ntab, 1), np.int32)
        tail = np.zeros(max(ntab, 1), np.float64)
        iexp = np.zeros(max(ntab, 1), np.int32)
        nke = np.zeros(max(ntab, 1), np.int32)
        for i, t in enumerate(self.tables):
            ut[i, :, : t.ngrid + 1] = t.coef.T
            ngrid[i] = t.ngrid
            tail[i] = t.tail
            iexp[i] = t.iexp
            if t.vk is not None:
                nke[i] = len(t.vk)
                ut[i, 0, t.ngrid + 1: t.ngrid + 1 + len(t.vk)] = t.vk
                ut[i, 0, t.ngrid + 1 + len(t.vk)] = t.vk0
        kvec = np.ascontiguousarray(self.kvec, dtype=np.float64).reshape(-1, 2)
        if kvec.shape[0] == 0:
            kvec = np.zeros((1, 2))
        s = QmcTables()
        s.ndim = 2
        s.ntypes = self.ntypes
        for i in range(self.ntypes):
            s.np[i] = self.np[i]
            s.hbs2m[i] = self.hbs2m[i]
        s.el[0], s.el[1] = self.el
        s.drt = self.drt
        s.ntab = ntab
        s.ldut = ldut
        s.ut = dptr(ut)
        s.ngrid = iptr(ngrid)
        s.tail = dptr(tail)
        s.iexp = iptr(iexp)
        s.nk_ewald = iptr(nke)
        for k in range(4):
            s.iu2[k] = int(self.iu2[k])
            s.iu3[k] = int(self.iu3[k])
            s.iub[k] = int(self.iub[k])
            s.iv2[k] = int(self.iv2[k])
            s.v2value[k] = 1.0
        s.v0 = self.v0
        s.nk = int(self.kvec.shape[0])
        s.kvec = dptr(kvec)
        s.irhok[0], s.irhok[1] = self.irhok
        s.nrhok = self.nrhok
        s.ntheta = self.ntheta
        s.adrift = self.adrift
        s.delta = 0.0
        return s, [ut, ngrid, tail, iexp, nke, kvec]


def read_deck(dirpath, runid=None):
    """Read runid, seed and the .sy deck in `dirpath` (input, reptation.f:12-710)."""
    if runid is None:
        with open(os.path.join(dirpath, "runid")) as f:
            runid = f.readline().split()[0]
    seed = (0, 0, 0, 1)
    sp = os.path.join(dirpath, "seed")
    if os.path.exists(sp):
        with open(sp) as f:
            w = f.read().split()
        if len(w) >= 4:
            seed = tuple(int(v) for v in w[:4])
    d = Deck(dirpath=dirpath, runid=runid, seed=seed)
    with open(os.path.join(dirpath, runid + ".sy")) as f:
        recs = read_words(f.readlines())
    for w in recs:
        if not w:
            continue
        key = w[0]
        if key == "end":
            break
        if key == "ndim":
            if int(w[1]) != 2:
                raise ValueError("only ndim 2 is supported")
        elif key == "v0":
            d.v0 = _ffloat(w[1])
        elif key == "adrift":
            d.adrift = _ffloat(w[1])
        elif key in _DRIVER_KEYWORDS:
            d.driver[key] = w[1:]
        elif key == "seed":
            d.seed = tuple(int(v) for v in w[1:5])
        elif key == "pbc":
            d.el = (_ffloat(w[1]), _ffloat(w[2]))
        elif key == "kspace":
            d.kvec = read_kvec(os.path.join(dirpath, w[1]))
        elif key == "rhok":
            d.nrhok += 1
            d.irhok[d._type(w[1])] = d.nrhok
        elif key == "type":
            d.ntypes += 1
            d.typenames.append(w[1])
            d.np.append(int(w[2]))
            d.hbs2m.append(_ffloat(w[3]))
            d.x_file.append(w[4] if len(w) > 4 else "")
        elif key in ("u2", "u3", "backflow", "v2"):
            it, jt = d._type(w[1]), d._type(w[2])
            fname = w[4] if key == "backflow" else w[3]
            i = d._table(fname)
            arr = {"u2": d.iu2, "u3": d.iu3, "backflow": d.iub, "v2": d.iv2}[key]
            arr[it * 2 + jt] = i
            arr[jt * 2 + it] = i
            if key == "v2":
                d.tables[i - 1].tail = _ffloat(w[4])
                d.tables[i - 1].iexp = int(w[5])
        elif key == "ntwists":
            d.ntheta = int(w[1])
        elif key in _UNSUPPORTED:
            raise ValueError(f"deck keyword '{key}' is outside the evaluation path")
        else:
            raise ValueError(f"{key} not a keyword")
    return d


def read_configs(paths, counts, ncfg=None):
    """Read list-directed configuration files, one per type, `counts[it]`
    lines of 2 reals per configuration (read_conf, reptation.f:2474-2485).
    Returns [nconf][nptot][2]."""
    per = []
    for p, n in zip(paths, counts):
        with open(p) as f:
            v = np.array([_ffloat(t) for t in f.read().split()], dtype=np.float64)
        per.append(v.reshape(-1, n, 2))
    nc = min(a.shape[0] for a in per)
    if ncfg is not None:
        nc = min(nc, ncfg)
    return np.ascontiguousarray(np.concatenate([a[:nc] for a in per], axis=1))
