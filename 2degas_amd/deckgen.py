"""Deck generator for the 2D electron gas in the reference's file formats.

Writes the files `input` reads (reptation.f:12-710): runid, seed,
<runid>.sy, the spline tables .u2/.u3/.b/.v (t_read, reptation.f:4689-4738)
and the k-vector list .k (kread, reptation.f:4740-4769).

Units follow the reference: lengths in rs, energies in Rydberg,
hbar^2/2m = 1/rs^2 and e^2 = 2/rs (setup.f:34-42, 452).  The box is square,
L = sqrt(pi N).

* k-vectors: the half space of the reciprocal lattice (first nonzero index
  positive), stably sorted by |k|^2, as krlv builds it
  (reptation.f:4962-5000); the first `nshell` shells are kept.
* Spline tables: values and slopes on the uniform grid r_i = i*drt, turned
  into the scaled cubic of `spline` (reptation.f:5223-5245): t1 = f,
  t2 = f' drt, t3 = 3(f1-f0) - (t2_1 + 2 t2_0), t4 = -2(f1-f0) + (t2_1+t2_0),
  last row zero.
* Ewald: the standard erfc split of e2/r in 2D.  Real space: table of
  r v_sr(r) = e2 erfc(alpha r) read with iexp = 1; k space:
  vk = (2 pi e2 / (A k)) erfc(k / 2 alpha) on the half-space list; the
  self and background constants go into v0 and vk0 = 0.  This is a
  textbook Ewald sum in place of the reference's optimized breakup fit
  (setup.f:287-542); both give the same Coulomb energy to their truncation
  errors.
"""
import math
import os

import numpy as np
from scipy.special import erfc


def klist(L, nshell):
    """Half-space reciprocal vectors of a square box, first nshell shells."""
    b = 2.0 * math.pi / L
    nmax = int(math.sqrt(nshell) * 2) + 8
    vecs = []
    # generation order of krlv: ix(1) fastest, ix(2) slowest
    for i2 in range(-nmax, nmax + 1):
        for i1 in range(-nmax, nmax + 1):
            nz = i1 if i1 != 0 else i2
            if nz <= 0:
                continue
            vecs.append((i1 * i1 + i2 * i2, i1, i2))
    # stable sort by |k|^2 (insertion before first strictly larger)
    vecs.sort(key=lambda v: v[0])
    shells = sorted(set(v[0] for v in vecs))[:nshell]
    last = shells[-1]
    if last > nmax * nmax:
        raise ValueError("increase nmax")
    out = [(i1 * b, i2 * b) for n2, i1, i2 in vecs if n2 <= last]
    return np.array(out, dtype=np.float64)


def spline_rows(f, df, drt, zero_last=True):
    """Scaled cubic rows of `spline` (reptation.f:5229-5243)."""
    nt = len(f) - 1
    t = np.zeros((nt + 1, 4))
    t[:, 0] = f
    t[:, 1] = df * drt
    for i in range(nt):
        d = t[i + 1, 0] - t[i, 0]
        t[i, 2] = 3.0 * d - (t[i + 1, 1] + 2.0 * t[i, 1])
        t[i, 3] = -2.0 * d + (t[i + 1, 1] + t[i, 1])
    if zero_last:
        t[nt] = 0.0
    return t


def write_table(path, rows, drt, routine="deckgen", params=(), kspace=None):
    nt = rows.shape[0] - 1
    with open(path, "w") as fh:
        fh.write(f" {nt} {float(drt)!r}\n")
        for r in rows:
            fh.write(" " + " ".join(f"{v: .17e}" for v in r) + "\n")
        fh.write(f" {routine}\n {len(params)}\n")
        for p in params:
            fh.write(f" {float(p)!r} 0\n")
        if kspace is not None:
            vk, vk0 = kspace
            fh.write(" kspace\n")
            fh.write(f" {len(vk)}\n")
            for v in vk:
                fh.write(f" {float(v)!r}\n")
            fh.write(f" {float(vk0)!r}\n")


def _cut(r, R):
    """Smooth cutoff (1 - r/R)^3 and its derivative, zero beyond R."""
    x = np.clip(1.0 - r / R, 0.0, None)
    return x ** 3, -3.0 * x ** 2 / R


def ewald_constants(N, L, e2, alpha, kv):
    """k-space coefficients vk (half-space list, so sum_half vk |rho_k|^2 is
    (1/2A) sum_{k != 0} v(k) |rho_k|^2) and v0 = self term
    -N alpha e2/sqrt(pi) plus the neutralizing background
    -N^2 e2 sqrt(pi) / (alpha A)."""
    A = L * L
    k = np.sqrt((kv ** 2).sum(1))
    vk = 2.0 * math.pi * e2 / (A * k) * erfc(k / (2.0 * alpha))
    self_e = -N * alpha * e2 / math.sqrt(math.pi)
    bg = -N * N * e2 * math.sqrt(math.pi) / (alpha * A)
    return vk, self_e + bg


def make_deck(dirpath, runid, nup, ndown, rs, nshell_ewald, nshell_k=None, ngrid=200,
              ngrid_u3=100, u3=True, backflow=True, eta_amp=0.2, ntwists=0, seed=(3535, 3467, 3262, 1549),
              param_seed=1, perturb=0.0, alpha=None):
    """Write a complete deck; returns a dict of the physical parameters."""
    os.makedirs(dirpath, exist_ok=True)
    rng = np.random.default_rng(param_seed)
    pert = lambda: 1.0 + perturb * rng.uniform(-1.0, 1.0)
    N = nup + ndown
    L = math.sqrt(math.pi * N)
    hbs2m = 1.0 / rs ** 2
    e2 = 2.0 / rs
    R = L / 2.0
    drt = R / ngrid
    r = np.arange(ngrid + 1) * drt
    # two-body: RPA-like u(r) = A (1 - exp(-r/F)) / r, cut smoothly at L/2
    Fc = math.sqrt(rs) * pert()
    Aamp = math.sqrt(rs) * 0.8 * pert()
    rr = np.where(r > 0, r, 1.0)
    ex = np.exp(-rr / Fc)
    base = np.where(r > 0, Aamp * (1.0 - ex) / rr, Aamp / Fc)
    dbase = np.where(r > 0, Aamp * (ex / (Fc * rr) - (1.0 - ex) / rr ** 2), -Aamp / (2 * Fc * Fc))
    c, dc = _cut(r, R)
    u2f = base * c
    u2d = dbase * c + base * dc
    # three-body xi(r): smooth seeded Gaussians on ngrid_u3 points (the
    # table ends at ngrid_u3*drt; beyond, the last zero row is used)
    ru = r[: ngrid_u3 + 1]
    Ru = ngrid_u3 * drt
    cu, dcu = _cut(ru, Ru)
    coef = rng.normal(0.0, 0.02, size=3)
    wid = np.array([0.6, 1.2, 2.0]) * max(1.0, Ru / 6.0)
    g = sum(cf * np.exp(-(ru / w) ** 2) for cf, w in zip(coef, wid))
    dg = sum(cf * np.exp(-(ru / w) ** 2) * (-2.0 * ru / w ** 2) for cf, w in zip(coef, wid))
    u3f = g * cu
    u3d = dg * cu + g * dcu
    # backflow eta(r) = a exp(-(r - 1)^2) in rs units (r0 = w = rs Bohr)
    a = eta_amp * pert()
    e = a * np.exp(-(r - 1.0) ** 2)
    de = e * (-2.0 * (r - 1.0))
    etaf = e * c
    etad = de * c + e * dc
    # Ewald
    kv_e = klist(L, nshell_ewald)
    kv = klist(L, nshell_k or nshell_ewald)
    if alpha is None:
        kmax = float(np.sqrt((kv_e ** 2).sum(1)).max())
        alpha = kmax / (2.0 * 5.6)
    vk, v0 = ewald_constants(N, L, e2, alpha, kv_e)
    vtab = e2 * erfc(alpha * r)
    dvtab = -e2 * 2.0 * alpha / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2)
    write_table(os.path.join(dirpath, runid + ".u2"), spline_rows(u2f, u2d, drt), drt, "rpa", [Aamp, Fc])
    if u3:
        write_table(os.path.join(dirpath, runid + ".u3"), spline_rows(u3f, u3d, drt), drt, "gauss3", list(coef))
    write_table(os.path.join(dirpath, runid + ".b"), spline_rows(etaf if backflow else 0 * etaf,
                                                              etad if backflow else 0 * etad, drt),
                drt, "eta_gauss", [a])
    write_table(os.path.join(dirpath, runid + ".v"), spline_rows(vtab, dvtab, drt), drt, "ewald_erfc",
                [alpha], kspace=(vk, 0.0))
    with open(os.path.join(dirpath, runid + ".k"), "w") as fh:
        for k in kv:
            fh.write(f" {float(k[0])!r} {float(k[1])!r}\n")
    types = [("up", nup)] + ([("down", ndown)] if ndown > 0 or nup > 0 else [])
    lines = ["ndim 2"]
    for name, n in types:
        lines.append(f"type {name} {n} {float(hbs2m)!r} {runid}.{name}.x")
    pairs = [("up", "up"), ("up", "down"), ("down", "down")]
    for kind, fn in (("u2", runid + ".u2"), ("u3", runid + ".u3")):
        if kind == "u3" and not u3:
            continue
        for p in pairs:
            lines.append(f"{kind} {p[0]} {p[1]} {fn}")
    for name, _ in types:
        lines.append(f"rhok {name}")
    lines.append(f"pbc {float(L)!r} {float(L)!r}")
    lines.append(f"kspace {runid}.k")
    for p in pairs:
        lines.append(f"backflow {p[0]} {p[1]} 1 {runid}.b")
    if ntwists:
        lines.append(f"ntwists {ntwists}")
    for p in pairs:
        lines.append(f"v2 {p[0]} {p[1]} {runid}.v 0. 1 1.")
    lines.append(f"v0 {float(v0)!r}")
    lines.append("end")
    for ln in lines:
        if len(ln) > 80:
            raise ValueError("deck record longer than 80 characters: " + ln)
    with open(os.path.join(dirpath, runid + ".sy"), "w") as fh:
        fh.write("\n".join(ln.ljust(80) for ln in lines) + "\n")
    with open(os.path.join(dirpath, "runid"), "w") as fh:
        fh.write(runid + "\n")
    with open(os.path.join(dirpath, "seed"), "w") as fh:
        fh.write(" ".join(str(s) for s in seed) + "\n")
    return dict(N=N, L=L, rs=rs, hbs2m=hbs2m, e2=e2, alpha=float(alpha), v0=float(v0), nk=len(kv),
                nk_ewald=len(kv_e), drt=drt)


def random_configs(nconf, nptot, L, seed):
    """Positions uniform in [-L/2, L/2)^2 from a seeded generator."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-L / 2, L / 2, size=(nconf, nptot, 2))


def lattice_configs(nconf, nptot, L, seed, jitter=0.15):
    """Positions on a jittered square grid: each configuration takes nptot of
    the ceil(sqrt(nptot))^2 cells in random order and displaces them by a
    Gaussian of width jitter * cell."""
    n = int(math.ceil(math.sqrt(nptot)))
    m = int(math.ceil(nptot / n))
    gx = (np.arange(n) + 0.5) / n * L - L / 2
    gy = (np.arange(m) + 0.5) / m * L - L / 2
    sites = np.array([(a, b) for a in gx for b in gy])
    rng = np.random.default_rng(seed)
    x = np.empty((nconf, nptot, 2))
    for c in range(nconf):
        x[c] = sites[rng.permutation(len(sites))[:nptot]] + rng.normal(0, jitter * L / n, size=(nptot, 2))
    return x


def write_configs(dirpath, runid, x, counts, names=("up", "down"), suffix=".0"):
    """Configuration files in read_conf's layout (reptation.f:2474-2485)."""
    off = 0
    for name, n in zip(names, counts):
        with open(os.path.join(dirpath, f"{runid}.{name}.x{suffix}"), "w") as fh:
            for c in x:
                for p in c[off:off + n]:
                    fh.write(f" {float(p[0])!r} {float(p[1])!r}\n")
        off += n
