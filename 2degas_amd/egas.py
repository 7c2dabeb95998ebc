"""The reference's deck generator, restated in numpy.

The reference builds its input tables with the program `egas` (setup.f) and
the table generators it shares with reptation.f.  This module restates the
parts a 2D electron gas deck needs, in the reference's own arithmetic order:

* `tgen` + `spline` (reptation.f:5007-5035, 5223-5245): a radial function
  and its derivative on r_i = i*drt, turned into the scaled cubic rows
  t1 = f, t2 = f' drt, t3, t4 that `getf` evaluates; the last row zeroed
  when the generator asks for it (spline_flag).
* the generators `u2_ob_cos` (reptation.f:5037-5111, the two-body
  pseudopotential of the shipped n26.u2), `u2_reatto` (reptation.f:
  6517-6562, n26.u3), `eta_yk` (reptation.f:9474-9512, the backflow of
  n26.b), `bozzo` (reptation.f:5177-5196, a Gaussian) and the smooth
  continuation `obsmooth` (reptation.f:6564-6589) they share.
* the optimized Ewald breakup of `ocpset` (setup.f:287-542): the k shells
  (`shells`, `fillk`, setup.f:859-936, 424-451), the bare 2D Coulomb
  coefficients 2 pi e2/(A k), the constrained least-squares fit of the
  long-range part by a polynomial times (r/a - 1)^3 (`fitpn`, setup.f:
  1033-1195, with `plint2`, `pmult`, `polyd`, `stfun` and the LINPACK
  Cholesky `spoco`/`spofa`/`sposl`, setup.f:1558-1601, 1611-1729,
  1730-1749, 1889-1901, 2121-2147), the Madelung check `mdlng`
  (setup.f:996-1032, with `gam`/`gammi`, setup.f:1479-1532), and
  `traduci` (setup.f:2227-2314): the real-space table with iexp = 1, the
  k-space coefficients of the first nlamb shells, and v0.

Where the reference uses a library routine the restatement calls the same
function from scipy: `mmbsjn` (setup.f:1902-2219, Bessel J_n by Miller's
recurrence) becomes scipy.special.jv.  Every constant keeps the value the
reference writes: ocpset, sang, fillk and fitpn use pi = 3.14159265358979,
mdlng pi = 3.1415926535, the generators acos(-1).
"""
import math

import numpy as np
from scipy.special import jv

PI_S = 3.14159265358979      # setup.f:318, 580, 438, 1043
PI_MDLNG = 3.1415926535      # setup.f:1000


# ---------------------------------------------------------------------------
# radial tables (reptation.f)

def spline(f, df, drt, spline_flag=1):
    """Rows (t1, t2, t3, t4) of `spline` (reptation.f:5223-5245)."""
    nt = len(f) - 1
    t = np.zeros((nt + 1, 4))
    t[:, 0] = f
    t[:, 1] = np.asarray(df) * drt
    for i in range(nt):
        t[i, 2] = 3.0 * (t[i + 1, 0] - t[i, 0]) - (t[i + 1, 1] + 2.0 * t[i, 1])
        t[i, 3] = -2.0 * (t[i + 1, 0] - t[i, 0]) + (t[i + 1, 1] + t[i, 1])
    if spline_flag:
        t[nt] = 0.0
    return t


def _obsmooth(r0, drt, i0, nt, f, df, c, xi, g, dg, ddg):
    """obsmooth (reptation.f:6564-6589): a quartic in r - c from c on, with
    the value, slope and curvature of the generator at c and zero slope and
    curvature at c + xi; returns beta0."""
    beta0 = -0.5 * dg * xi - ddg * xi ** 2 / 12.0
    beta1 = dg
    beta2 = 0.5 * ddg
    beta3 = -dg / xi ** 2 - 2.0 / 3.0 * ddg / xi
    beta4 = 0.5 * dg / xi ** 3 + 0.25 * ddg / xi ** 2
    for i in range(i0, nt + 1):
        r = r0 + i * drt
        d = r - c
        f[i] = beta0 + beta1 * d + beta2 * d ** 2 + beta3 * d ** 3 + beta4 * d ** 4
        df[i] = beta1 + 2.0 * beta2 * d + 3.0 * beta3 * d ** 2 + 4.0 * beta4 * d ** 3
    return beta0


def u2_ob_cos(p, nt, drt, r0=0.0):
    """u2_ob_cos (reptation.f:5037-5111): [(a r + b r^2)/(1 + c^2 r + d^2 r^2)
    + e] * [1 + sum_k e_k (1 - cos(k_k r))], smoothly continued from
    x = drt*(nt*5/6) by obsmooth and shifted to vanish at the table end.
    p = (a, b, c, d, e, e_1, ..., e_n)."""
    p = [float(v) for v in p]
    n = len(p) - 5
    p3 = p[2] ** 2
    p4 = p[3] ** 2
    x = drt * (nt * 5 // 6)
    xi = drt * nt - x
    pl = math.acos(-1.0) / (2.0 * drt * nt)
    c, dc, ddc = 1.0, 0.0, 0.0
    for k in range(1, n + 1):
        kk = 2 * k * pl
        c = c + p[4 + k] * (1 - math.cos(kk * x))
        dc = dc + p[4 + k] * math.sin(kk * x) * kk
        ddc = ddc + p[4 + k] * math.cos(kk * x) * kk ** 2
    den = 1 + p3 * x + p4 * x ** 2
    g = (p[0] * x + p[1] * x ** 2) / den + p[4]
    dg = (p[0] + 2 * p[1] * x) / den - (p[0] * x + p[1] * x ** 2) * (p3 + 2 * p4 * x) / den ** 2
    ddg = (2 * p[1] / den
           - 2 * (p[0] + 2 * p[1] * x) * (p3 + 2 * p4 * x) / den ** 2
           - (p[0] * x + p[1] * x ** 2) * 2 * p4 / den ** 2
           + 2 * (p[0] * x + p[1] * x ** 2) * (p3 + 2 * p4 * x) ** 2 / den ** 3)
    ddg = ddg + ddc + 2 * dg * dc
    dg = g * dc + c * dg
    g = g * c
    f = np.zeros(nt + 1)
    df = np.zeros(nt + 1)
    i = 0
    while i <= nt:
        r = max(r0 + i * drt, 1e-10)
        if not r < x:
            break
        cr, dcr = 1.0, 0.0
        for k in range(1, n + 1):
            kk = 2 * k * pl
            cr = cr + p[4 + k] * (1 - math.cos(kk * r))
            dcr = dcr + p[4 + k] * math.sin(kk * r) * kk
        dr = 1 + p3 * r + p4 * r ** 2
        fi = (p[0] * r + p[1] * r ** 2) / dr + p[4]
        dfi = (p[0] + 2 * p[1] * r) / dr - (p[0] * r + p[1] * r ** 2) * (p3 + 2 * p4 * r) / dr ** 2
        df[i] = dfi * cr + fi * dcr
        f[i] = fi * cr
        i += 1
    beta0 = _obsmooth(r0, drt, i, nt, f, df, x, xi, g, dg, ddg)
    f[:i] = f[:i] - g + beta0
    return f, df, 1


def u2_reatto(p, nt, drt, r0=0.0):
    """u2_reatto (reptation.f:6517-6562): a/r^b + c exp(-d (r-e)^2)
    + f exp(-g (r-h)^2), smoothly continued by obsmooth; p = (a..h)."""
    p = [float(v) for v in p]
    x = drt * (nt * 5 // 6)
    xi = drt * nt - x
    e1 = math.exp(-p[3] * (x - p[4]) ** 2)
    e2 = math.exp(-p[6] * (x - p[7]) ** 2)
    g = p[0] / x ** p[1] + p[2] * e1 + p[5] * e2
    dg = (-p[1] * p[0] / x ** (p[1] + 1.0) - 2 * p[2] * p[3] * (x - p[4]) * e1
          - 2 * p[5] * p[6] * (x - p[7]) * e2)
    ddg = ((p[1] + 1.0) * p[1] * p[0] / x ** (p[1] + 2.0)
           - (2 * p[2] * p[3] - 4 * p[2] * p[3] ** 2 * (x - p[4]) ** 2) * e1
           - (2 * p[5] * p[6] - 4 * p[5] * p[6] ** 2 * (x - p[7]) ** 2) * e2)
    f = np.zeros(nt + 1)
    df = np.zeros(nt + 1)
    i = 0
    while i <= nt:
        r = max(r0 + i * drt, 1e-10)
        if not r < x:
            break
        a1 = math.exp(-p[3] * (r - p[4]) ** 2)
        a2 = math.exp(-p[6] * (r - p[7]) ** 2)
        f[i] = p[0] / r ** p[1] + p[2] * a1 + p[5] * a2
        df[i] = (-p[1] * p[0] / r ** (p[1] + 1.0) - 2 * p[2] * p[3] * (r - p[4]) * a1
                 - 2 * p[5] * p[6] * (r - p[7]) * a2)
        i += 1
    beta0 = _obsmooth(r0, drt, i, nt, f, df, x, xi, g, dg, ddg)
    f[:i] = f[:i] - g + beta0
    return f, df, 1


def eta_yk(p, nt, drt, r0=0.0):
    """eta_yk (reptation.f:9474-9512), the backflow of Kwon et al. 1998:
    lambda (1 + s r)/(r_B + w r + r^3.5), continued by obsmooth;
    p = (lambda, s, r_B, w)."""
    p = [float(v) for v in p]
    p[2] = math.sqrt(p[2] ** 2)
    c = drt * (nt * 5 // 6)
    xi = drt * nt - c
    g = p[0] * (1 + p[1] * c) / (p[2] + p[3] * c + c ** 3.5)
    dg = (p[1] * g - (p[3] + 3.5 * c ** 2.5) * g ** 2 / p[0]) / (1 + p[1] * c)
    ddg = -(12 * c ** 2 * g ** 2 + 2 * g * dg * (p[3] + 3.5 * c ** 2.5)) / (1 + p[1] * c) / p[0]
    f = np.zeros(nt + 1)
    df = np.zeros(nt + 1)
    i = 0
    while i <= nt:
        r = r0 + i * drt
        if not r < c:
            break
        f[i] = p[0] * (1 + p[1] * r) / (p[2] + p[3] * r + r ** 3.5)
        df[i] = (p[1] * f[i] - (p[3] + 3.5 * r ** 2.5) * f[i] ** 2 / p[0]) / (1 + p[1] * r)
        i += 1
    beta0 = _obsmooth(r0, drt, i, nt, f, df, c, xi, g, dg, ddg)
    f[:i] = f[:i] - g + beta0
    return f, df, 1


def bozzo(p, nt, drt, r0=0.0):
    """bozzo (reptation.f:5177-5196): a exp(-b (r - c)^2), last row kept."""
    a, b, c = (float(v) for v in p)
    r = r0 + np.arange(nt + 1) * drt
    f = a * np.exp(-b * (r - c) ** 2)
    return f, -2 * b * (r - c) * f, 0


GENERATORS = {"u2_ob_cos": u2_ob_cos, "u2_reatto": u2_reatto, "eta_yk": eta_yk, "bozzo": bozzo}


def tgen(name, params, nt, drt):
    """tgen (reptation.f:5007-5035): the spline rows of generator `name`."""
    f, df, flag = GENERATORS[name](params, nt, drt)
    return spline(f, df, drt, flag)


def format_d20_12(v):
    """One value in the reference's table edit descriptor 4d20.12
    (reptation.f:5024): 0.dddddddddddd in the mantissa, a signed two-digit
    exponent, right-justified in 20 characters."""
    if v == 0.0:
        return "  0.000000000000D+00"
    m, e = f"{v:.11e}".split("e")
    e = int(e) + 1
    digits = m.replace("-", "").replace(".", "")
    s = ("-" if v < 0 else "") + "0." + digits + "D" + ("-" if e < 0 else "+") + f"{abs(e):02d}"
    return s.rjust(20)


def write_tgen_table(path, rows, drt, routine, params, flags=None):
    """A table file as tgen writes it: `nt drt`, the rows in 4d20.12, the
    routine name, the parameter count and one `value flag` line each
    (reptation.f:5018-5033); flags default to 1 (optimizable)."""
    nt = rows.shape[0] - 1
    flags = flags if flags is not None else [1] * len(params)
    with open(path, "w") as fh:
        fh.write(f" {nt} {float(drt)!r}\n")
        for r in rows:
            fh.write("".join(format_d20_12(float(v)) for v in r) + "\n")
        fh.write(f" {routine}\n {len(params)}\n")
        for v, fl in zip(params, flags):
            fh.write(f" {float(v)!r} {int(fl)}\n")


# ---------------------------------------------------------------------------
# Ewald breakup (setup.f)

def sang(ndim):
    """Solid angle in ndim dimensions (setup.f:576-583)."""
    return 2 * PI_S * (ndim - 1)


def sites_box(natoms, vol, ndim=2):
    """The box side `sites` computes (setup.f:584-716): the lattice with
    the fewest vacancies among 1 and 2 sites per square cell, ncell cells a
    side, side ncell * (vol/ncell^2)^(1/2)."""
    small = 1.0e-3
    nvac, nxtal = natoms, 0
    for l in range(1, ndim + 1):
        npuc = 2 ** (l - 1)
        nc = int((natoms / npuc) ** (1.0 / ndim) + 1.0 - small)
        nvact = npuc * nc ** ndim - natoms
        if nvact < nvac:
            nvac, nxtal = nvact, npuc
    npuc = nxtal
    xcell = (float(natoms) / float(npuc)) ** (1.0 / float(ndim))
    nc = int(xcell / 1.0 + 1.0 - small)
    npts = nc ** ndim
    a = (vol / npts) ** (1.0 / ndim)
    return nc * a * 1.0, npuc, npts * npuc - natoms


def ggrid_linear(n, r0, r1):
    """ggrid 'LINEAR' (setup.f:828-848), 1-based r(1..n) as a 0-based array."""
    dr = (r1 - r0) / (n - 1)
    return np.array([r0 + i * dr for i in range(n)])


def shells(ndim, a, cut):
    """shells (setup.f:859-936): the half-space vectors with |k| <= cut,
    ordered by norm as its insertion sort leaves them; returns the shell
    norms and the running vector counts kmult (1-based lists)."""
    c2 = cut ** 2
    nkspan = [int(0.00001 + abs(cut / a[l])) for l in range(ndim)]
    icount = [-nkspan[l] for l in range(ndim)]
    npts = 1
    for l in range(ndim):
        npts *= 2 * nkspan[l] + 1
    vecs = []
    for _ in range(npts):
        rsq = 0.0
        nzero = 0
        x = [0.0] * ndim
        keep = True
        for l in range(ndim):
            if nzero == 0:
                nzero = icount[l]
            x[l] = icount[l] * a[l]
            rsq = rsq + x[l] ** 2
            if rsq > c2:
                keep = False
                break
        if keep and nzero > 0:
            kj = len(vecs)
            for j, v in enumerate(vecs):
                rks = 0.0
                for l in range(ndim):
                    rks = rks + v[l] ** 2
                if rks >= rsq:
                    kj = j
                    break
            vecs.insert(kj, tuple(x))
        for l in range(ndim):
            icount[l] += 1
            if icount[l] <= nkspan[l]:
                break
            icount[l] = -nkspan[l]
    rknorm, kmult = [], []
    rnow = 0.0
    for i, v in enumerate(vecs):
        rsq = 0.0
        for l in range(ndim):
            rsq = rsq + v[l] ** 2
        if rsq - rnow > 0.001 * rnow:
            rknorm.append(0.0)
            kmult.append(0)
        rnow = rsq
        rknorm[-1] = math.sqrt(rnow)
        kmult[-1] = i + 1
    return rknorm, kmult


def fillk(shell_norms, kmult, argek, ndim, vol):
    """fillk (setup.f:424-451): rknorm(1) = 0 and the shells from
    rknorm(2), their weights (vectors counted with their negatives), and a
    uniform continuation up to argek weighted by the volume of k space.
    Returns 1-based arrays (index 0 unused) rknorm, wtk and nshex1."""
    nshlls = len(shell_norms)
    rknorm = {1: 0.0}
    for s, v in enumerate(shell_norms):
        rknorm[s + 2] = v
    wtk = {1: 1.0, 2: float(2 * kmult[0])}
    for k in range(2, nshlls + 1):
        wtk[k + 1] = float(2 * (kmult[k - 1] - kmult[k - 2]))
    dk = rknorm[2] * 0.25
    cutk = rknorm[nshlls]
    nshex = int(nshlls + (argek - cutk) / dk)
    con = vol * sang(ndim) / (ndim * (2 * PI_S) ** ndim)
    vdown = float(1 + 2 * kmult[nshlls - 1])
    for k in range(nshlls + 1, nshex + 1):
        vup = con * (cutk + dk * (k - nshlls)) ** ndim
        rknorm[k + 1] = cutk + dk * (k - nshlls - 0.5)
        wtk[k + 1] = vup - vdown
        vdown = vup
    nshex1 = nshex + 1
    rk = np.zeros(nshex1 + 1)
    wt = np.zeros(nshex1 + 1)
    for k in range(1, nshex1 + 1):
        rk[k] = rknorm[k]
        wt[k] = wtk[k]
    return rk, wt, nshex1


def gam(a):
    """Gamma function by the rational approximation of `gam`
    (setup.f:1479-1505)."""
    p = (.2514886580600251e+05, .5298348466186016e+04, .6177655268060726e+04,
         .2509926126029017e+03, .5785107455981657e+03, -.2316113056472773e+02,
         .2021018352970918e+02)
    q = (.1000000000000000e+01, .2513925177055871e+05, .1985721823627555e+05,
         -.7373357770095750e+04, -.5047154372852621e+03, .3632214014257158e+03,
         -.3119624946361091e+02)
    z = a
    fctr = 1.0
    while z < 1.0:
        fctr = fctr / z
        z = z + 1.0
    while z > 2.0:
        z = z - 1.0
        fctr = fctr * z
    tnum = (((((p[6] * z + p[5]) * z + p[4]) * z + p[3]) * z + p[2]) * z + p[1]) * z + p[0]
    tden = (((((q[6] * z + q[5]) * z + q[4]) * z + q[3]) * z + q[2]) * z + q[1]) * z + q[0]
    return tnum / tden * fctr


def gammi(a, x):
    """Incomplete gamma G(a, x) and G(a, 0) of `gammi` (setup.f:1506-1532):
    series below x = 1.2, a 20-term continued fraction above."""
    y = abs(x)
    xa = y ** a
    gm = gam(a)
    if y <= 1.2:
        gami = gm - xa / a
        term = -xa
        for n in range(1, 16):
            term = -term * y / n
            gami = gami + term / (a + n)
        return gami, gm
    term = 2 * y
    m = 20
    for _ in range(20):
        term = y + (m - a) * term / (term + m)
        m -= 1
    return math.exp(-y) * xa / term, gm


def mdlng(ndim, p, ell, e2):
    """Madelung energy of an r^-p interaction on the periodic lattice of
    `ell` by an Ewald sum of incomplete gamma functions (setup.f:996-1032)."""
    pi = PI_MDLNG
    eps = 1.0e-9
    limit = int(1 + math.sqrt(-math.log(eps) / pi))
    vol = 1.0
    for l in range(ndim):
        vol = vol * ell[l]
    icount = [-limit] * ndim
    alpha = pi * vol ** (-2.0 / ndim)
    v = (-2 * alpha ** (.5 * p) / (p * gam(.5 * p))
         - 2 * pi ** (.5 * ndim) / ((ndim - p) * vol * gam(.5 * p) * alpha ** (.5 * (ndim - p))))
    for _ in range((2 * limit + 1) ** ndim):
        r2 = 0.0
        rk2 = 0.0
        for l in range(ndim):
            r2 = r2 + (icount[l] * ell[l]) ** 2
            rk2 = rk2 + (2 * pi * icount[l] / ell[l]) ** 2
        if r2 > 0:
            p2 = .5 * p
            gr, gr0 = gammi(p2, alpha * r2)
            gk, _ = gammi(.5 * (ndim - p), rk2 / (4.0 * alpha))
            v = v + pi ** (.5 * ndim) * (4.0 / rk2) ** (.5 * (ndim - p)) * gk / (vol * gr0) \
                + gr / (gr0 * r2 ** p2)
        for l in range(ndim):
            icount[l] += 1
            if icount[l] <= limit:
                break
            icount[l] = -limit
    return v * e2


def pmult(c, n, np_, pn):
    """pmult (setup.f:1889-1901): differences for the factor (x-1)^np, then
    the scale pn; c is 1-based (index 0 unused)."""
    for k in range(1, np_ + 1):
        for i in range(1, n + np_ - k + 1):
            c[i] = c[i + 1] - c[i]
    for i in range(1, n + 1):
        c[i] = pn * c[i]


def plint2(r, n, i0, pn, np_):
    """plint2 (setup.f:1558-1601): integrals of J_0(r x) x^(i+i0-1) (x-1)^np
    over [0, 1] for i = 1..n by Abramowitz-Stegun 11.1.1, times pn; the
    Bessel values come from scipy.special.jv (the reference's mmbsjn
    computes the same J_n).  Returns a 1-based array."""
    c = np.zeros(n + np_ + 2)
    if r >= 1.0e-5:
        nbf = min(max(100, 6 * int(r)), 999)
        bess = np.concatenate([[0.0], jv(np.arange(nbf), r)])  # bess[k] = J_{k-1}
        fact = 1.0e-10
        con = 2.0 / (math.exp(1.0) * r)
    for i in range(1, n + np_ + 1):
        nn = i + i0 - 2
        if r < 1.0e-5:
            ci = 1.0
        else:
            ci = bess[2]
            rat = 1.0
            for k in range(4, nbf + 1, 2):
                rat = rat * float(-nn + k - 4) / float(nn + k)
                term = rat * bess[k] * (k - 1)
                ci = ci + term
                if abs(rat) < (con * k) ** k * fact:
                    break
            ci = 2.0 * ci / r
        c[i] = ci / (i + i0)
    pmult(c, n, np_, pn)
    return c


def _ddot(n, x, y):
    """LINPACK ddot with unit strides (reptation.f:6087-6121): the first
    mod(n,5) products one at a time, then five per step, left to right."""
    s = 0.0
    m = n % 5
    for i in range(m):
        s = s + x[i] * y[i]
    for i in range(m, n, 5):
        s = s + x[i] * y[i] + x[i + 1] * y[i + 1] + x[i + 2] * y[i + 2] + x[i + 3] * y[i + 3] \
            + x[i + 4] * y[i + 4]
    return s


def _spofa(A, n):
    """LINPACK spofa (setup.f:2121-2147): Cholesky A = R^T R in the upper
    triangle, A[row][col] 0-based."""
    for j in range(n):
        s = 0.0
        for k in range(j):
            t = A[k][j] - _ddot(k, [A[i][k] for i in range(k)], [A[i][j] for i in range(k)])
            t = t / A[k][k]
            A[k][j] = t
            s = s + t * t
        s = A[j][j] - s
        if s <= 0.0:
            raise ValueError(f"fitpn: matrix not positive definite (info={j + 1})")
        A[j][j] = math.sqrt(s)


def _sposl(A, n, b):
    """LINPACK sposl (setup.f:1713-1729): solve R^T R x = b in place."""
    for k in range(n):
        t = _ddot(k, [A[i][k] for i in range(k)], b[:k])
        b[k] = (b[k] - t) / A[k][k]
    for kb in range(1, n + 1):
        k = n - kb
        b[k] = b[k] / A[k][k]
        t = -b[k]
        for i in range(k):
            b[i] = b[i] + t * A[i][k]


def fitpn(n, v, rk, wt, nk, nf, b0, i0, vol, r, ndim=2, ifcon=1):
    """fitpn (setup.f:1033-1195): fit the Fourier components v(k) for k
    beyond shell nf by a polynomial of m = n - 3 terms times (r/a - 1)^3 on
    [0, a], a = r(lptable), with the cusp constraint; subtract the fit from
    every v(k).  Arrays are 1-based (index 0 unused); v is updated in place.
    Returns (poly b[1..m], rms, the self term vmad, the table of the fitted
    short-range part and its derivative on the grid r)."""
    a = r[-1]
    pn = 2 * PI_S * (ndim - 1) * a ** ndim / vol
    np_ = 3
    m = n - np_
    if ifcon != 1:
        mp, i2, beta1, beta2 = m, 1, 0.0, 0.0
    else:
        mp, i2 = m - 1, 2
        anp = 1 - 2 * (np_ % 2)
        if i0 != 0:
            beta1 = -anp * b0 * a / float(np_)
            beta2 = 1.0 / float(np_)
        else:
            beta1 = anp * b0 / a
            beta2 = 0.0
    # hs(j, i) for j, i in 1..m (lhs), column m+1 the rhs
    lhs = [[0.0] * (m + 1) for _ in range(m + 1)]
    rhs = [0.0] * (m + 1)
    chisq = 0.0
    for k in range(nf + 2, nk + 1):
        c = plint2(rk[k] * a, m, i0, pn, np_)
        vv = v[k] - beta1 * c[1]
        chisq = chisq + wt[k] * vv ** 2
        c[2] = c[2] + beta2 * c[1]
        for i in range(i2, m + 1):
            rhs[i] = rhs[i] + wt[k] * vv * c[i]
            for j in range(i2, m + 1):
                lhs[j][i] = lhs[j][i] + wt[k] * c[i] * c[j]
    A = [[lhs[i2 + rr][i2 + cc] for cc in range(mp)] for rr in range(mp)]
    _spofa(A, mp)
    bcopy = [rhs[i] for i in range(i2, m + 1)]
    x = list(bcopy)
    _sposl(A, mp, x)
    sol = [0.0] * (m + 1)
    for idx, i in enumerate(range(i2, m + 1)):
        sol[i] = x[idx]
        chisq = chisq - bcopy[idx] * x[idx]
    rms = math.sqrt(chisq) if chisq > 0 else float("nan")
    if ifcon == 1:
        sol[1] = beta1 + beta2 * sol[2]
    for k in range(1, nk + 1):
        c = plint2(rk[k] * a, m, i0, pn, np_)
        for i in range(1, m + 1):
            v[k] = v[k] - sol[i] * c[i]
    b = np.array(sol[1:m + 1])
    vmad = (-1.0) ** np_ * (b[1] - np_ * b[0])
    # polyd and the table (setup.f:1135-1157)
    ap = np.zeros(m)
    for i in range(2, m + 1):
        ap[i - 2] = (i - 1.0) * b[i - 1]
    iexp = 1 - i0
    ai = a ** iexp
    v1t = np.zeros(len(r))
    v2t = np.zeros(len(r))
    for idx, ri in enumerate(r):
        xx = ri / a
        v1 = _stfun(xx, b)
        v2 = _stfun(xx, ap)
        if iexp != 0:
            v1 = v1 * ai
            v2 = v2 * ai
        v2 = v2 * (xx - 1.0) ** np_ + np_ * (xx - 1.0) ** (np_ - 1) * v1
        v1 = v1 * (xx - 1.0) ** np_
        v1t[idx] = v1
        v2t[idx] = v2 / a
    return dict(poly=b, rms=rms, vmad_self=vmad, table=v1t, dtable=v2t, pn=pn, a=a,
                beta1=beta1, m=m)


def _stfun(x, a):
    """Horner evaluation of `stfun` (setup.f:1740-1749)."""
    s = 0.0
    for i in range(len(a)):
        s = a[len(a) - i - 1] + x * s
    return s


def ocpset(nparts, rs, nt, nlamb=18, cutk=8.0, argek=14.0):
    """The Coulomb part of ocpset (setup.f:287-542) followed by traduci
    (setup.f:2227-2314): the optimized breakup of e2/r, e2 = 2/rs, in a
    square box of nparts electrons at density 1/pi (lengths in rs).

    Returns the real-space table values vr and slopes dvr on the ggrid mesh
    (501 points for nt = 500 from 1e-6 to L/2), the k-space coefficients of
    the half-space vectors of the first nlamb shells (vk, each shell value
    repeated over its vectors), v0 per particle (traduci's (v(1) np +
    self)/2), and the diagnostics te.os prints: the ideal and Ewald-summed
    Madelung energies, the fit's rms error, the polynomial and the fitted
    Fourier components of the first nlamb + 1 shells.  The reference fixes
    nlamb = 18, cutk = 8 and argek = 14 (setup.f:431, 414-417); more shells
    need the k cutoffs raised with them."""
    ndim = 2
    sangle = sang(ndim)
    ro = ndim / sangle
    vol = nparts / ro
    ell, npuc, nvac = sites_box(nparts, vol, ndim)
    lptable = nt + 1
    cutr = .5 * ell
    mpoly = 10
    eps = 1.0e-6
    rv = ggrid_linear(lptable, eps, cutr)
    tpiell = [2 * PI_S / ell] * ndim
    norms, kmult = shells(ndim, tpiell, cutk)
    rknorm, wtk, nshex1 = fillk(norms, kmult, argek, ndim, vol)
    e2 = 2.0 / rs
    vkbare = np.zeros(nshex1 + 1)
    for k in range(2, nshex1 + 1):
        vkbare[k] = (e2 * sangle) / (vol * rknorm[k] ** (ndim - 1))
    vmad_ideal = -3.90026492 * e2 / vol ** (1.0 / 2.0)
    vmad_ewald = mdlng(ndim, 1.0, [ell] * ndim, e2)
    v = vkbare.copy()
    fit = fitpn(mpoly, v, rknorm, wtk, nshex1, nlamb, e2, 0, vol, rv, ndim, 1)
    computed = fit["vmad_self"]
    for k in range(1, nshex1 + 1):
        computed = computed + wtk[k] * v[k]
    # traduci: the k-space coefficients of shells 2..nlamb+1, each repeated
    # over its wt/2 half-space vectors, and v0
    vk = []
    for k in range(2, nlamb + 2):
        mult = int(.5 * wtk[k] + 1.0e-4)
        vk.extend([v[k]] * mult)
    v0 = (v[1] * nparts + fit["vmad_self"]) * 0.5
    return dict(vr=fit["table"], dvr=fit["dtable"], vk=np.array(vk), v0=v0, ell=ell, vol=vol,
                rv=rv, vmad_ideal=vmad_ideal, vmad_ewald=vmad_ewald, vmad_fit=computed,
                rms=fit["rms"], poly=fit["poly"], k=rknorm[1:nlamb + 2].copy(),
                f=v[1:nlamb + 2].copy(), nk=nshex1, rknorm=rknorm, wtk=wtk, vkbare=vkbare,
                beta1=fit["beta1"], pn=fit["pn"], a=fit["a"], npuc=npuc, nvac=nvac)
