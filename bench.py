"""Throughput of batched local-energy evaluation on one MI355X.

One step = one evaluation (compute_properties(1) for every walker) of a
batch of walkers whose positions are already resident in HBM.  Workload:
N=122 (61+61) electrons, rs=10, Slater-Jastrow-backflow with u2 + u3,
Ewald over the first 200 k-shells (900 k-vectors), 4096 walkers per launch,
synthetic positions (a jittered 12x12 grid per walker) and a generated deck.

Prints one JSON line: the throughput (value, ms_per_step), the workload
(config), the dominant kernel's roofline (algorithmic TFLOP/s against the
FP64 matrix-core peak, HBM bytes per launch from the committed counter
summary), the per-kernel HIP-event means (kernel_ms) and the reference's
CPU rate on this host (cpu_baseline).  --rqmc and --serve measure the
reptation and served-job workloads instead.
"""
import argparse
import importlib
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6


KERNELS = ["pairs", "det", "fgrad", "cross", "final"]


def gj_path():
    """The library's determinant path: k_gjm unless QMC_NO_GJ is set."""
    return os.environ.get("QMC_NO_GJ") is None


def kernel_flops(nup, ndown, nk_ewald):
    """Algorithmic flops per walker of each kernel (pairs, determinant,
    drift/diagonal terms, cross, final).  The determinant is k_gjm (matrix
    cores; k_fgrad then does the diagonal terms) unless QMC_NO_GJ selects
    k_det, which does both; the determinant's count is the Gauss-Jordan
    solve's, whichever kernel runs it.  The cross term is counted as the
    work of its formulation: per spin the symmetric product S = V V^T of
    the 2m x 2N backflow Jacobian, 2m (2m+1) 2N flops, and the contraction
    with M, 2 (2m)^2.  (The per-evaluation formula qmc_flops_per_eval counts
    24 N m^2 for the cross term, the reference's own loops; that count is
    above what any kernel here executes, so it is reported apart, as the
    formula rate.)"""
    n = nup + ndown
    pairs = 160.0 * n * (n - 1) / 2 + 5.0 * n * nk_ewald + 20.0 * nk_ewald
    gj = gj_path()
    det = sum(24.0 * m ** 3 + (0.0 if gj else 37.0 * m * m) for m in (nup, ndown))
    fgrad = sum(37.0 * m * m for m in (nup, ndown)) if gj else 0.0
    cross = sum(2.0 * m * (2 * m + 1) * 2.0 * n + 2.0 * (2 * m) ** 2 for m in (nup, ndown) if m > 0)
    return [pairs, det, fgrad, cross, 0.0]


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary (profiles/traffic_fp64_n122_w4096.md, copy-calibrated),
    or None when that summary has no row for it."""
    path = os.path.join(ROOT, "profiles", "traffic_fp64_n122_w4096.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == kernel:
            return float(cells[5]) * 1e6
    return None


def cpu_baseline(deckdir, x, seconds_target=12.0):
    """The reference Fortran (oracle/_ref/refeval, timing mode) run as one
    independent process per host core on a bounded sample."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if not pyoracle.have_ref():
        return None
    # one process per core this process may run on (its affinity mask)
    cores = len(os.sched_getaffinity(0))
    # calibrate: one process, a few evaluations
    work = tempfile.mkdtemp(prefix="degas_cpu_")
    raw, _ = pyoracle.run_refeval(deckdir, os.path.join(work, "cal"), 3, x[:64], nsteps=4)
    t1 = float(np.frombuffer(raw[20:28], dtype=np.float64)[0]) / 4
    nev = max(4, int(seconds_target / max(t1, 1e-6)))
    procs = []
    for p in range(cores):
        wd = os.path.join(work, f"p{p}")
        os.makedirs(wd, exist_ok=True)
        for fn in os.listdir(deckdir):
            src = os.path.join(deckdir, fn)
            if os.path.isfile(src):
                with open(src, "rb") as a, open(os.path.join(wd, fn), "wb") as b:
                    b.write(a.read())
        np.ascontiguousarray(x[(p * 64) % len(x):(p * 64) % len(x) + 64]).tofile(os.path.join(wd, "cfg.bin"))
        with open(os.path.join(wd, "refeval.ctl"), "w") as f:
            f.write(f"3\n1 1\n{nev}\n0.0\ncfg.bin\nout.bin\n")
    t0 = time.time()
    for p in range(cores):
        wd = os.path.join(work, f"p{p}")
        procs.append(subprocess.Popen([os.path.join(pyoracle.REF_DIR, "refeval")], cwd=wd,
                                      stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
    for pr in procs:
        pr.wait()
    wall = time.time() - t0
    total = cores * nev
    with open(os.path.join(pyoracle.REF_DIR, "COMPILER")) as f:
        comp = f.read().strip()
    return {"value": total / wall, "unit": "evals/s", "cores": cores, "kind": "reference",
            "host_cpu_count": os.cpu_count(),
            "sample": f"{nev} compute_properties(1) per process x {cores} processes (the cores "
                      f"of this process's affinity mask; the host has {os.cpu_count()}), "
                      f"same deck and synthetic walkers, {wall:.1f} s wall ({comp})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nw", type=int, default=4096)
    ap.add_argument("--nup", type=int, default=61)
    ap.add_argument("--ndown", type=int, default=61)
    ap.add_argument("--rs", type=float, default=10.0)
    ap.add_argument("--shells", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the single-electron and served-jobs legs of the default line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the size-based phase chunk of the pair kernels instead of qmc_autotune's")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--rqmc", action="store_true",
                    help="reptation workload instead: N=58 (29+29) rs=5 backflow, 512 reptiles "
                         "x 200 beads, delta 0.01 Ry^-1 (tau 0.02 Ha^-1); --steps moves are timed")
    ap.add_argument("--reptiles", type=int, default=512)
    ap.add_argument("--ntau", type=int, default=200)
    ap.add_argument("--mdelta", type=int, default=4)
    ap.add_argument("--serve", action="store_true",
                    help="served workload instead: Poisson arrivals of independent VMC jobs "
                         "(N in {26,58,74,122} or polarized {37,49,61}, rs in [1,20], +-10% "
                         "parameters) run to 1e-4 Ha (2e-4 Ry) per electron")
    ap.add_argument("--vmc", action="store_true",
                    help="VMC steps of the plain N=26 workload instead: 13+13, rs=5, u2 only, "
                         "zero backflow, 60 k-shells, 256 walkers; the step replayed as a "
                         "captured hipGraph and launched directly (QMC_NO_GRAPH)")
    ap.add_argument("--dmc", action="store_true",
                    help="DMC steps/s with fixed-population branching (N=122, 4096 walkers; N=26, 256)")
    ap.add_argument("--sem", action="store_true",
                    help="single-electron moves with Sherman-Morrison updates instead: plain "
                         "Slater-Jastrow N=122 (61+61, rs=10, Pade u2, zero backflow, 200 k-shells) "
                         "with 4096 walkers and N=26 with 256, FP64 and mixed inverses")
    ap.add_argument("--serve-moves", choices=("all", "single", "compare"), default="all",
                    help="--serve: the reference's all-electron moves (jobs with the n26 "
                         "backflow), packed single-electron moves (jobs without backflow), or "
                         "both on the same backflow-free job mix")
    ap.add_argument("--jobs", type=int, default=40)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--max-blocks", type=int, default=1000,
                    help="--serve: blocks after which a job stops without its target")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device-to-device copies: the known-size "
                         "kernel that scales the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every walker batch is independent, and this build "
              "runs on a single MI355X (--gpus 1).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "copies": 3}))
        return
    hip = importlib.import_module("2degas_amd.hip")
    deck = importlib.import_module("2degas_amd.deck")
    dg = importlib.import_module("2degas_amd.deckgen")
    if args.rqmc:
        return bench_rqmc(args, hip, deck, dg)
    if args.vmc:
        return bench_vmc(args, hip, deck, dg)
    if args.dmc:
        return bench_dmc(args, hip, deck, dg)
    if args.sem:
        return bench_sem(args, hip, deck, dg)
    if args.serve:
        return bench_serve(args)
    tmp = tempfile.mkdtemp(prefix="degas_bench_")
    deckdir = os.path.join(tmp, "deck")
    info = dg.make_deck(deckdir, "bench", args.nup, args.ndown, args.rs, args.shells)
    d = deck.read_deck(deckdir)
    N = d.nptot
    x = dg.lattice_configs(args.nw, N, d.el[0], 12345)
    ctx = hip.Context(d, twist=(1, 1), device=0)
    tune = None
    if not args.no_autotune:
        (k64, kmix), (ms64, msmix) = ctx.autotune(args.nw)
        tune = {"kchunk_fp64": k64, "kchunk_mixed": kmix, "tuning_ms_fp64": ms64, "tuning_ms_mixed": msmix}
    dev = torch.device("cuda:0")
    xt = torch.tensor(x, dtype=torch.float64, device=dev)
    props = torch.zeros((args.nw, ctx.info["n_props"]), dtype=torch.float64, device=dev)
    g = torch.zeros((args.nw, N, 2), dtype=torch.float64, device=dev)
    h = torch.zeros((args.nw, 2), dtype=torch.float64, device=dev)
    g2 = torch.zeros(args.nw, dtype=torch.float64, device=dev)
    st = torch.zeros(args.nw, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ctx.eval_device(xt, props, g, h, g2, st, stream=stream)
    torch.cuda.synchronize()
    ctx.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.eval_device(xt, props, g, h, g2, st, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_k, nl = ctx.timing_read()
    ctx.timing(False)
    dt = t1 - t0
    evals = args.nw * args.steps
    value = evals / dt
    guard = int((st.cpu().numpy() & 1).sum())
    ref64 = {"props": props.cpu().numpy(), "g": g.cpu().numpy()}
    mixed = bench_mixed(args, ctx, xt, props, g, h, g2, st, stream, ref64, value)
    kf = kernel_flops(args.nup, args.ndown, info["nk_ewald"])
    names = KERNELS
    gj = gj_path()
    part = any(0 < m <= 48 for m in (args.nup, args.ndown))
    det_kernel = ("k_gjm_part" if part else "k_gjm") if gj else "k_det"
    kname = {"pairs": "k_pairs", "det": det_kernel, "fgrad": "k_fgrad",
             "cross": "k_cross", "final": "k_final"}
    dom = int(np.argmax(ms_k))
    achieved = kf[dom] * args.nw / (ms_k[dom] * 1e-3) / 1e12
    total_flops = hip.flops_per_eval(args.nup, args.ndown, info["nk_ewald"])
    out = {
        "metric": "local-energy evals/s at N=122, mixed vs FP64 GPU; p50/p99 job latency",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"N={N} ({args.nup}+{args.ndown}) rs={args.rs} Slater-Jastrow-backflow "
                               f"u2+u3, Ewald {args.shells} k-shells ({info['nk_ewald']} k), "
                               f"{args.nw} walkers per launch, FP64 path",
                   "walkers": args.nw, "N": N, "nk_ewald": info["nk_ewald"], "precision": "fp64"},
        "roofline": {"bound": "mfma", "kernel": kname[names[dom]], "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": pmc_traffic(kname[names[dom]])},
        "kernel_ms": dict(zip(names, [float(v) for v in ms_k])),
        "kernel_frac_fp64_peak": {n: (kf[i] * args.nw / (ms_k[i] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
                                      if ms_k[i] > 0 else None) for i, n in enumerate(names)},
        "eval_tflops": sum(kf) * value / 1e12,
        "eval_frac_fp64_peak": sum(kf) * value / 1e12 / FP64_PEAK_TFLOPS,
        "eval_tflops_formula": total_flops * value / 1e12,
        "guard_fired": guard,
        "mixed": mixed,
        "autotune": tune,
    }
    ctx.close()
    if not args.no_extras:
        out["single_electron"] = single_electron_leg(hip, deck, dg)
        out["serve"] = serve_leg(args)
    if not args.no_cpu:
        cb = cpu_baseline(deckdir, x, args.cpu_seconds)
        out["cpu_baseline"] = cb
    print(json.dumps(out))


def single_electron_leg(hip, deck, dg):
    """Beside the evaluation rate: single-electron sweeps at N=122 (61+61,
    rs=10, Pade u2, zero backflow, 200 k-shells, 4096 walkers), FP64 and
    with FP32 inverses, the FP64 refresh, and one full evaluation of the
    same walkers, event-timed."""
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    tmp = tempfile.mkdtemp(prefix="degas_sel_")
    dg.make_deck(os.path.join(tmp, "deck"), "n122", 61, 61, 10.0, 200, u3=False, backflow="zero", jastrow="pade")
    d = deck.read_deck(os.path.join(tmp, "deck"))
    nw = 4096
    ctx = hip.Context(d, twist=(1, 1), device=0)
    x = dg.lattice_configs(nw, d.nptot, d.el[0], 99)
    sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
    wk = hip.Walkers(ctx, x, sts)
    wk.start()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(n):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / n

    res = {"workload": "N=122 (61+61) rs=10, Pade u2, zero backflow, 200 k-shells, 4096 walkers, step 0.6"}
    for prec, name in ((abi.QMC_PREC_FP64, "fp64"), (abi.QMC_PREC_MIXED, "fp32_inverse")):
        sm = hip.SingleMoves(wk, precision=prec)
        sm.refresh()
        ms = timed(lambda: sm.sweep(0.6), 10)
        res[name] = {"sweep_ms": ms, "electron_moves_per_s": 1e3 / ms * nw * d.nptot,
                     "refresh_ms": timed(sm.refresh, 3)}
        sm = None
    res["full_evaluation_ms"] = timed(wk.start, 5)
    wk = None
    ctx.close()
    return res


def serve_leg(args):
    """The served-jobs half of the metric: 40 VMC jobs with Poisson
    arrivals at 20/s (the --serve workload), p50/p99 time to the error bar."""
    sv = importlib.import_module("2degas_amd.serve")
    specs = sv.poisson_jobs(40, 20.0, seed=2024, nw=256, target=2e-4, max_blocks=1000, steps_per_block=20)
    _, summary = sv.serve(specs)
    return {"workload": "40 VMC jobs, Poisson 20/s, N in 26/58/74/122 or polarized 37/49/61, rs U[1,20], "
                        "Pade u2 + n26 backflow +-10%, 256 walkers, target 1e-4 Ha/electron",
            **{k: summary[k] for k in ("p50_latency_s", "p99_latency_s", "jobs_per_s", "converged", "capped",
                                       "wall_s")}}


def bench_mixed(args, ctx, xt, props, g, h, g2, st, stream, ref64, value64):
    """The mixed-precision leg on the same batch: QMC_PREC_MIXED (FP32 table
    evaluations of u2, u3 and the pair potential, FP32 rho_k phases and
    chunk sums; FP64 geometry, backflow, determinant, accumulators), timed
    as the FP64 leg, and its error against the FP64 evaluation of the same
    walkers, per walker: E_L against |epot| + |ekin| + |efp|, -log psi, and
    the drift as max_i |g_i - g_i^64| / max_i |g_i^64|."""
    import numpy as np
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    pm = torch.zeros_like(props)
    gm = torch.zeros_like(g)
    for _ in range(args.warmup):
        ctx.eval_device(xt, pm, gm, h, g2, st, precision=abi.QMC_PREC_MIXED, stream=stream)
    torch.cuda.synchronize()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.eval_device(xt, pm, gm, h, g2, st, precision=abi.QMC_PREC_MIXED, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_k, _ = ctx.timing_read()
    ctx.timing(False)
    v = args.nw * args.steps / dt
    info = ctx.info
    p64, pmx = ref64["props"], pm.cpu().numpy()
    j = lambda k: info[k] - 1
    es = np.abs(p64[:, j("jpot0")]) + np.abs(p64[:, j("jkin0")]) + np.abs(p64[:, j("jefp")])
    e_el = np.abs(pmx[:, j("jetot")] - p64[:, j("jetot")]) / np.maximum(np.abs(p64[:, j("jetot")]), es)
    e_ltf = np.abs(pmx[:, j("jltf")] - p64[:, j("jltf")]) / np.abs(p64[:, j("jltf")])
    g64, gmx = ref64["g"], gm.cpu().numpy()
    e_dr = np.max(np.abs(gmx - g64).reshape(len(g64), -1), axis=1) / \
        np.max(np.abs(g64).reshape(len(g64), -1), axis=1)
    q = lambda a: {"max": float(np.max(a)), "median": float(np.median(a))}
    return {"value": v, "unit": "evals/s", "ms_per_step": dt / args.steps * 1e3,
            "ratio_to_fp64": v / value64, "kernel_ms": dict(zip(KERNELS, [float(x) for x in ms_k])),
            "error_vs_fp64": {"elocal": q(e_el), "ltf": q(e_ltf), "drift": q(e_dr)},
            "target_3x_met": bool(v / value64 >= 3.0),
            "target_1e-6_met": bool(np.max(e_el) <= 1e-6)}


def bench_vmc(args, hip, deck, dg):
    """VMC steps/s (move + evaluation + Metropolis test + block sums of every
    walker) on the plain N=26 workload with 256 walkers, the step captured
    as a hipGraph and replayed, beside the same steps launched directly."""
    import numpy as np
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    nw = 256
    tmp = tempfile.mkdtemp(prefix="degas_vmc_")
    deckdir = os.path.join(tmp, "deck")
    dg.make_deck(deckdir, "c26", 13, 13, 5.0, 60, u3=False, backflow="zero")
    d = deck.read_deck(deckdir)
    x = dg.lattice_configs(nw, d.nptot, d.el[0], 4321)
    sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
    res = {}
    for mode in ("graph", "direct"):
        if mode == "direct":
            os.environ["QMC_NO_GRAPH"] = "1"
        try:
            ctx = hip.Context(d, twist=(1, 1), device=0)
        finally:
            os.environ.pop("QMC_NO_GRAPH", None)
        wk = hip.Walkers(ctx, x, sts)
        wk.start(delta=0.2)
        wk.sweep(max(args.warmup, 2), 0.2)
        torch.cuda.synchronize()
        nsteps = 200
        t0 = time.perf_counter()
        for _ in range(args.steps):
            wk.sweep(nsteps, 0.2)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[mode] = args.steps * nsteps / dt
        acc = float((wk.age == 0).double().mean())
        wk = None
        ctx.close()
    out = {
        "metric": "VMC steps/s (plain N=26, 256 walkers)",
        "value": res["graph"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps * 200,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / res["graph"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "VMC N=26 (13+13) rs=5 u2 only, zero backflow, 60 k-shells (232 k), "
                               "256 walkers, delta 0.2", "walkers": nw},
        "walker_moves_per_s": res["graph"] * nw,
        "direct_launch_steps_per_s": res["direct"],
        "graph_speedup": res["graph"] / res["direct"],
        "acceptance_last_step": acc,
    }
    print(json.dumps(out))


def bench_dmc(args, hip, deck, dg):
    """DMC steps/s (a move of every walker, its weight, the weighted block
    sums, the branching and the copies: qmc_dmc_sweep with mmstep 1) on the
    N=122 workload with 4096 walkers and on the plain N=26 one with 256,
    and the time of the branching alone (one thread over the walkers)."""
    import numpy as np
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    res = {}
    for tag, nu, rs, shells, nw, kw in (("n122", 61, 10.0, 200, 4096, {}),
                                        ("n26", 13, 5.0, 60, 256, dict(u3=False, backflow="zero"))):
        tmp = tempfile.mkdtemp(prefix="degas_dmc_")
        deckdir = os.path.join(tmp, "deck")
        dg.make_deck(deckdir, tag, nu, nu, rs, shells, **kw)
        d = deck.read_deck(deckdir)
        x = dg.lattice_configs(nw, d.nptot, d.el[0], 777)
        sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
        ctx = hip.Context(d, twist=(1, 1), device=0)
        wk = hip.Walkers(ctx, x, sts)
        delta = 0.05 * rs * rs / d.nptot
        wk.start(delta=delta)
        wk.sweep(20, delta)
        torch.cuda.synchronize()
        etrial = float(wk.props[:, ctx.info["jetot"] - 1].mean().item())
        dm = hip.Dmc(wk, [3535, 3467, 3262, 1549])
        dm.sweep(max(args.warmup, 2), delta, etrial)
        torch.cuda.synchronize()
        nsteps = args.steps if tag == "n122" else 10 * args.steps
        t0 = time.perf_counter()
        dm.sweep(nsteps, delta, etrial)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        copies = int((dm.src != torch.arange(nw, device=dm.src.device, dtype=torch.int32)).sum().item())
        # the branching alone, on the last step's weights
        w0 = torch.rand(nw, dtype=torch.float64, device=wk.x.device) + 0.5
        wt, r2 = w0.clone(), dm.rng2.clone()
        s0, c0 = dm.src.clone(), dm.counts.clone()
        hip.dmc_branch(ctx, wt, r2, s0, c0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(20):
            wt.copy_(w0)
            hip.dmc_branch(ctx, wt, r2, s0, c0)
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t1) / 20
        res[tag] = {"steps_per_s": nsteps / dt, "walker_moves_per_s": nsteps * nw / dt, "walkers": nw,
                    "N": d.nptot, "ms_per_step": dt / nsteps * 1e3, "branch_ms": tb * 1e3,
                    "copies_last_step": copies}
        wk = dm = None
        ctx.close()
    out = {
        "metric": "DMC steps/s (N=122, 4096 walkers; move, weights, branching, copies)",
        "value": res["n122"]["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["n122"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "DMC with fixed-population branching, N=122 (61+61) rs=10 "
                               "u2+u3+backflow, 200 k-shells, 4096 walkers, mmstep 1"},
        "workloads": res,
    }
    print(json.dumps(out))


def bench_sem(args, hip, deck, dg):
    """Single-electron sweeps/s (qmc_sem_moves over every electron, one spin
    per launch) with FP64 and FP32 inverses, beside the inverse refresh, the
    full evaluation that gives a sweep's local energy, and one all-electron
    VMC step on the same deck and walkers; then N=58 with u2 + u3 at rs 1, 5
    and 10 (8192 walkers), where the FP32 inverses' refresh interval K is
    tuned against the measured drift and the log|psi| carried over K moves
    is checked against full evaluations."""
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    res = {}
    for tag, nu, rs, shells, nw, sigma in (("n122", 61, 10.0, 200, 4096, 0.6), ("n26", 13, 5.0, 60, 256, 0.5)):
        tmp = tempfile.mkdtemp(prefix="degas_sem_")
        deckdir = os.path.join(tmp, "deck")
        dg.make_deck(deckdir, tag, nu, nu, rs, shells, u3=False, backflow="zero", jastrow="pade")
        d = deck.read_deck(deckdir)
        x = dg.lattice_configs(nw, d.nptot, d.el[0], 99)
        sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
        ctx = hip.Context(d, twist=(1, 1), device=0)
        wk = hip.Walkers(ctx, x, sts)
        wk.start()
        r = {"N": d.nptot, "walkers": nw, "sigma": sigma}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

        def timed(fn, n):
            fn()
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(n):
                fn()
            ev[1].record()
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / n

        nsw = max(args.steps, 5)
        for prec, name in ((abi.QMC_PREC_FP64, "fp64"), (abi.QMC_PREC_MIXED, "mixed")):
            sm = hip.SingleMoves(wk, precision=prec)
            tile_ms, tile = sm.tune_tile(sigma)
            sm.refresh()
            for _ in range(max(args.warmup, 2)):
                sm.sweep(sigma)
            sm.nacc.zero_()
            ms = timed(lambda: sm.sweep(sigma), nsw)
            acc = float(sm.nacc.sum().item()) / (nw * d.nptot * (nsw + 1))
            r[name] = {"sweep_ms": ms, "sweeps_per_s": 1e3 / ms,
                       "electron_moves_per_s": 1e3 / ms * nw * d.nptot,
                       "refresh_ms": timed(sm.refresh, 5), "acceptance": acc,
                       "tile_threads_per_row": tile, "tile_ms": tile_ms}
            sm = None
        r["eval_ms"] = timed(wk.start, 5)
        delta = 0.05 * rs * rs / d.nptot
        wk.start(delta=delta)
        r["all_electron_step_ms"] = timed(lambda: wk.sweep(1, delta), 10)
        res[tag] = r
        wk = None
        ctx.close()
    # the batched plain-vs-mixed shape: N=58 (29+29), u2 + u3, zero backflow,
    # 8192 walkers, step 0.5 rs, rs 1, 5, 10; FP32 inverses refreshed in FP64
    # every K moves, K tuned against the measured drift
    for rs in (1.0, 5.0, 10.0):
        tag = f"n58_rs{rs:g}"
        tmp = tempfile.mkdtemp(prefix="degas_sem_")
        deckdir = os.path.join(tmp, "deck")
        dg.make_deck(deckdir, "n58", 29, 29, rs, 60, u3=True, backflow="zero", jastrow="pade")
        d = deck.read_deck(deckdir)
        nw = 8192
        x = dg.lattice_configs(nw, d.nptot, d.el[0], 58)
        sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nw)]
        ctx = hip.Context(d, twist=(1, 1), device=0)
        wk = hip.Walkers(ctx, x, sts)
        wk.start()
        jltf = ctx.info["jltf"] - 1
        r = {"N": d.nptot, "walkers": nw, "rs": rs, "sigma": 0.5}
        for prec, name in ((abi.QMC_PREC_FP64, "fp64"), (abi.QMC_PREC_MIXED, "mixed")):
            sm = hip.SingleMoves(wk, precision=prec)
            tile_ms, tile = sm.tune_tile(0.5)
            sm.refresh()
            for _ in range(max(args.warmup, 5)):
                sm.sweep(0.5)
            sm.nacc.zero_()
            ms = timed(lambda: sm.sweep(0.5), max(args.steps, 5))
            acc = float(sm.nacc.sum().item()) / (nw * d.nptot * (max(args.steps, 5) + 1))
            rms = timed(sm.refresh, 3)
            kbest, drifts = sm.tune_refresh(0.5)
            # log|psi| carried over K moves against full evaluations
            sm.refresh()
            wk.start()
            ltf0 = wk.props[:, jltf].clone()
            sm.dlogpsi.zero_()
            sm.advance(kbest, 0.5)
            wk.start()
            torch.cuda.synchronize()
            lerr = float((sm.dlogpsi + (wk.props[:, jltf] - ltf0)).abs().max().item())
            r[name] = {"sweep_ms": ms, "refresh_ms": rms, "acceptance": acc, "K": kbest,
                       "tile_threads_per_row": tile, "tile_ms": tile_ms,
                       "drift_by_K": drifts, "logpsi_err_after_K": lerr,
                       "sweep_ms_with_refresh": ms + rms * d.nptot / kbest,
                       "electron_moves_per_s": 1e3 / ms * nw * d.nptot}
            sm = None
        r["eval_ms"] = timed(wk.start, 5)
        res[tag] = r
        wk = None
        ctx.close()
        # the same shape with the Gaussian backflow of the batched workload
        # (0.2 exp(-((r - rs)/rs)^2)): every move is a full evaluation
        tag = f"n58_rs{rs:g}_backflow"
        deckdir = os.path.join(tmp, "deck_bf")
        dg.make_deck(deckdir, "n58b", 29, 29, rs, 60, u3=True, backflow="gauss", jastrow="pade")
        d = deck.read_deck(deckdir)
        ctx = hip.Context(d, twist=(1, 1), device=0)
        wk = hip.Walkers(ctx, x, sts)
        wk.start()
        r = {"N": d.nptot, "walkers": nw, "rs": rs, "sigma": 0.5}
        for prec, name in ((abi.QMC_PREC_FP64, "fp64"), (abi.QMC_PREC_MIXED, "mixed")):
            sm = hip.SingleMoves(wk, precision=prec)
            sm.nacc.zero_()
            ms = timed(lambda: sm.sweep(0.5), 2)
            r[name] = {"sweep_ms": ms, "acceptance": float(sm.nacc.sum().item()) / (nw * d.nptot * 3),
                       "electron_moves_per_s": 1e3 / ms * nw * d.nptot}
            sm = None
        r["eval_ms"] = timed(wk.start, 3)
        res[tag] = r
        wk = None
        ctx.close()
    # packing: 16 small jobs (N 26, 42, 58, 74, four each; 256 walkers; u2
    # only, zero backflow), one sweep each as separate calls on one stream
    # against one packed call per spin and row class
    jobs = []
    for k in range(16):
        nu = (13, 21, 29, 37)[k % 4]
        tmp = tempfile.mkdtemp(prefix="degas_sem_")
        deckdir = os.path.join(tmp, "deck")
        dg.make_deck(deckdir, "pk", nu, nu, 2.0 + k % 5, 12, u3=False, backflow="zero", jastrow="pade")
        d = deck.read_deck(deckdir)
        ctx = hip.Context(d, twist=(1, 1), device=0)
        x = dg.lattice_configs(256, d.nptot, d.el[0], 700 + k)
        sts = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + q + 256 * k) + 1) % 4096]) for q in range(256)]
        wk = hip.Walkers(ctx, x, sts)
        wk.start()
        sm = hip.SingleMoves(wk)
        sm.refresh()
        jobs.append((sm, 0, d.nptot, 0.5))
    sep_ms = timed(lambda: [sm.moves(a, b, sg) for sm, a, b, sg in jobs], 10)
    pak_ms = timed(lambda: hip.sem_moves_packed(jobs), 10)
    res["packed16"] = {"jobs": 16, "walkers_per_job": 256, "N": [13 * 2, 21 * 2, 29 * 2, 37 * 2],
                       "separate_ms": sep_ms, "packed_ms": pak_ms, "speedup": sep_ms / pak_ms}
    jobs = None
    n = res["n122"]
    out = {
        "metric": "single-electron sweeps/s (N=122, 4096 walkers, FP64 inverse)",
        "value": n["fp64"]["sweeps_per_s"],
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": max(args.steps, 5),
        "warmup": args.warmup,
        "ms_per_step": n["fp64"]["sweep_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single-electron moves with Sherman-Morrison updates, plain "
                               "Slater-Jastrow N=122 (61+61) rs=10, Pade u2, zero backflow, "
                               "200 k-shells, 4096 walkers, Gaussian step 0.6"},
        "workloads": res,
    }
    print(json.dumps(out))


def bench_serve(args):
    """Continuous batching of independent VMC jobs: p50/p99 time from arrival
    to the requested error bar, jobs/s and aggregate local-energy evals/s."""
    sv = importlib.import_module("2degas_amd.serve")
    if args.serve_moves != "all":
        return bench_serve_single(args, sv)
    specs = sv.poisson_jobs(args.jobs, args.rate, seed=2024, nw=256, target=2e-4,
                            max_blocks=args.max_blocks, steps_per_block=20)
    results, summary = sv.serve(specs)
    out = {
        "metric": "served VMC jobs: p50/p99 time to 1e-4 Ha/electron",
        "value": summary["p99_latency_s"],
        "unit": "s (p99)",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.jobs} VMC jobs, Poisson {args.rate}/s, N in 26/58/74/122 "
                               "or polarized 37/49/61, rs U[1,20], params +-10%, 256 walkers, "
                               "20-step blocks",
                   "jobs": args.jobs, "rate": args.rate},
        "p50_latency_s": summary["p50_latency_s"],
        "p99_latency_s": summary["p99_latency_s"],
        "jobs_per_s": summary["jobs_per_s"],
        "evals_per_s": summary["evals_per_s"],
        "converged": summary["converged"],
        "capped": summary["capped"],
        "latency_over": "jobs that reached the target (capped jobs stopped at max_blocks)",
        "wall_s": summary["wall_s"],
        "max_blocks": args.max_blocks,
        "per_job": [{"job": r.job_id, "N": s.nup + s.ndown, "nup": s.nup, "rs": round(s.rs, 3),
                     "blocks": len(r.blocks), "error": r.error, "converged": r.converged,
                     "latency_s": r.latency, "device_bytes": r.device_bytes}
                    for s in specs for r in [results[s.job_id]]],
    }
    print(json.dumps(out))


def bench_serve_single(args, sv):
    """The served workload with jobs without backflow (u2 only): packed
    single-electron sweeps (20-sweep blocks, every electron moved once per
    sweep, the local energy evaluated after each sweep), and with
    --serve-moves compare the all-electron service on the same jobs."""
    modes = ("single", "all") if args.serve_moves == "compare" else ("single",)
    legs = {}
    for mode in modes:
        specs = sv.poisson_jobs(args.jobs, args.rate, seed=2024, nw=256, target=2e-4,
                                max_blocks=args.max_blocks, steps_per_block=20, backflow="zero")
        results, summary = sv.serve(specs, moves=mode)
        legs[mode] = dict(summary, per_job=[{"job": r.job_id, "N": s.nup + s.ndown, "rs": round(s.rs, 3),
                                             "blocks": len(r.blocks), "converged": r.converged,
                                             "latency_s": r.latency}
                                            for s in specs for r in [results[s.job_id]]])
    lead = legs["single"]
    out = {
        "metric": "served VMC jobs (no backflow), packed single-electron moves: p50/p99 time to 1e-4 Ha/electron",
        "value": lead["p99_latency_s"],
        "unit": "s (p99)",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.jobs} VMC jobs, Poisson {args.rate}/s, N in 26/58/74/122 "
                               "or polarized 37/49/61, rs U[1,20], Pade u2, zero backflow, 256 walkers, "
                               "20-sweep blocks", "jobs": args.jobs, "rate": args.rate},
        "legs": legs,
    }
    print(json.dumps(out))


def bench_rqmc(args, hip, deck, dg):
    """Reptation moves/s on the RQMC workload: one move = direction and
    ndelta drawn, mdelta batched bead evaluations of every reptile, path
    weight and reptation test (reptation_move + reptation_metropolis_test)."""
    import numpy as np
    import torch
    abi = importlib.import_module("2degas_amd.abi")
    tmp = tempfile.mkdtemp(prefix="degas_rqmc_")
    deckdir = os.path.join(tmp, "deck")
    info = dg.make_deck(deckdir, "rqmc", 29, 29, 5.0, 60)
    d = deck.read_deck(deckdir)
    N = d.nptot
    nr, ntau, md = args.reptiles, args.ntau, args.mdelta
    ctx = hip.Context(d, twist=(1, 1), device=0)
    # initial paths: a jittered lattice per reptile followed by small
    # Gaussian steps along the imaginary-time axis
    x0 = dg.lattice_configs(nr, N, d.el[0], 777)
    rng = np.random.default_rng(778)
    beads = x0[:, None] + np.cumsum(rng.normal(0.0, 0.02, size=(nr, ntau, N, 2)), axis=1)
    beads[:, 0] = x0
    # reptile k draws the stream of MPI rank k (seed(4) = 2*(seed(4)+k)+1 of
    # the shipped seed 3535 3467 3262 1549, reptation.f:37-39)
    states = [abi.limbs_to_state([3535, 3467, 3262, (2 * (1549 + k) + 1) % 4096]) for k in range(nr)]
    prm = hip.rmc_params(0.01, -0.3, 1e30, 1e30, ntau // 10, md)
    rep = hip.Reptiles(ctx, nr, ntau, md, states)
    t0 = time.perf_counter()
    rep.start(beads, prm)
    torch.cuda.synchronize()
    t_start = time.perf_counter() - t0
    rep.sweep(args.warmup, prm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep.sweep(args.steps, prm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = rep.status()
    moves = nr * args.steps
    out = {
        "metric": "reptation moves/s (RQMC, N=58)",
        "value": moves / dt,
        "unit": "moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"RQMC N={N} (29+29) rs=5 u2+u3+backflow, {nr} reptiles x {ntau} beads, "
                               f"mdelta {md}, delta 0.01 (tau 0.02 Ha^-1)",
                   "reptiles": nr, "ntau": ntau, "mdelta": md},
        # ndelta is uniform in 1..mdelta and only moving beads are evaluated
        "bead_evals_per_s_expected": nr * (md + 1) / 2 * args.steps / dt,
        "start_s": t_start,
        "status_nonzero": int(np.sum(st != 0)),
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
