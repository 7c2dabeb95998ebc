"""Continuous-batching service for independent VMC jobs on one GPU.

A job is a VMC run of its own system (N, rs, polarization, wavefunction
parameters perturbed from a seed) that ends when the statistical error of its
local energy per electron reaches the requested target, measured with the
reference's block error (averages(3), reptation.f:4324-4346:
err = sqrt(|<b^2> - <b>^2| / (n - 1)) over n block means b).

Jobs are admitted in arrival order while their device memory fits the
budget: a job's context is created first (its tables go to the device) and
the library reports, through qmc_ctx_info, what the context holds and what
each walker adds to its workspace and VMC buffers; with the walker state
that is exactly what the job allocates.  Admitted jobs keep their walkers
resident on the device, and run
concurrently, each on its own HIP stream: a job's next block of VMC sweeps
is queued as soon as its previous block has finished (polled through an
event), so small and large systems progress at their own pace while the
device overlaps their kernels.  Finished blocks are folded into each job's
statistics and jobs that reach their target retire, freeing memory for the
next arrivals.
A job's numbers do not depend on the mix it ran in: it owns its context,
walkers and random streams.
"""
import math
import os
import tempfile
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, deck, deckgen, hip


@dataclass
class JobSpec:
    job_id: int
    nup: int
    ndown: int
    rs: float
    param_seed: int = 1        # seed of the +-perturb parameter draw
    perturb: float = 0.1       # relative perturbation of the Jastrow/backflow amplitudes
    nw: int = 256              # walkers (reference MPI ranks)
    delta: float = 0.0         # time step; 0 picks rs^2/N (all-electron var = 2/N in rs units)
    steps_per_block: int = 20
    equil_blocks: int = 1      # blocks discarded before statistics start
    target: float = 2e-4       # error of E per electron (Ry) that ends the job
    min_blocks: int = 4
    max_blocks: int = 400
    shells: int = 20           # Ewald k-shells of the generated deck
    # the generated deck (deckgen.make_deck): a two-body Jastrow that
    # decreases to zero at L/2 for every N and rs (the n26 parameters turn
    # into an attractive well at larger boxes), no three-body term, the n26
    # backflow; the free parameters perturbed by `perturb`
    jastrow: str = "pade"
    u3: bool = False
    backflow: str = "yk"
    # consecutive rejections after which a move is accepted (the reference's
    # mage, metropolis_test): walkers started next to a node, where every
    # proposal is rejected, move again instead of freezing their energy
    mage: float = 50.0
    # single-electron moves (serve(moves="single")): the Gaussian step per
    # dimension, in units of rs; a sweep moves every electron once
    sigma: float = 0.5
    arrival: float = 0.0       # seconds after the service starts
    seed_limbs: tuple = (3535, 3467, 3262, 1549)


@dataclass
class JobResult:
    job_id: int
    energy: float              # mean local energy per electron (Ry)
    error: float               # block error of that mean
    blocks: list = field(default_factory=list)  # block means after equilibration
    arrival: float = 0.0
    admitted: float = 0.0
    done: float = 0.0
    evals: int = 0
    converged: bool = False
    predicted_bytes: int = 0   # device memory admission counted for the job
    device_bytes: int = 0      # what its context and walkers held at the end

    @property
    def latency(self):
        return self.done - self.arrival


def single_bytes(ctx, nw):
    """Bytes of a SingleMoves state for nw walkers (FP64 inverses of the
    larger spin, the log|psi| and count accumulators) and of the context's
    device copy of its system description."""
    mm = max(int(v) for v in ctx.deck.np[:ctx.deck.ntypes])
    return nw * (2 * mm * mm * 16 + 2 * 8)


def job_bytes(ctx, nw, single=False):
    """Device memory of a VMC job of nw walkers on context ctx: what the
    context holds (tables, g-vectors), its evaluation workspace and VMC
    proposal buffers for nw walkers (per-walker sizes from qmc_ctx_info) and
    the walker state tensors; with single-electron moves, their state."""
    m = ctx.device_bytes()
    return (m["dev_bytes"] + nw * (m["eval_bytes_per_walker"] + m["vmc_bytes_per_walker"])
            + hip.Walkers.state_bytes(nw, ctx.deck.nptot, ctx.info["n_props"])
            + (single_bytes(ctx, nw) if single else 0))


def block_error(blocks):
    """The reference's error of the mean of n block means
    (reptation.f:4336-4346)."""
    n = len(blocks)
    if n < 2:
        return math.inf
    b = np.asarray(blocks, dtype=np.float64)
    c1, c2 = b.sum() / n, (b * b).sum() / n
    return math.sqrt(abs(c2 - c1 * c1) / (n - 1))


class _Running:
    """A job: its deck and context from construction (the tables on the
    device, the memory it will need known), its walkers from start()."""

    def __init__(self, spec, workdir, single=False):
        self.spec = spec
        self.single = single
        if single and spec.backflow != "zero":
            raise ValueError(f"job {spec.job_id}: single-electron moves are packed only without backflow")
        ddir = os.path.join(workdir, f"job{spec.job_id}")
        deckgen.make_deck(ddir, "job", spec.nup, spec.ndown, spec.rs, spec.shells,
                          jastrow=spec.jastrow, u3=spec.u3, backflow=spec.backflow,
                          param_seed=spec.param_seed, perturb=spec.perturb)
        self.deck = deck.read_deck(ddir)
        self.ctx = hip.Context(self.deck, twist=(1, 1))
        self.jetot = self.ctx.info["jetot"] - 1
        self.need = job_bytes(self.ctx, spec.nw, single)

    def start(self, t_admit):
        import torch
        spec = self.spec
        self.t_admit = t_admit
        x = deckgen.lattice_configs(spec.nw, self.deck.nptot, self.deck.el[0],
                                    1000 + spec.param_seed)
        s = list(spec.seed_limbs)
        # walker k draws the stream of MPI rank k (reptation.f:37-39)
        states = [abi.limbs_to_state([s[0], s[1], s[2], (2 * (s[3] + k) + 1) % 4096])
                  for k in range(spec.nw)]
        # all-electron jobs run on streams of their own; single-electron
        # jobs share the service's stream, where their moves are packed
        self.stream = torch.cuda.current_stream() if self.single else torch.cuda.Stream()
        self.delta = spec.delta if spec.delta > 0 else spec.rs * spec.rs / self.deck.nptot
        with torch.cuda.stream(self.stream):
            self.walkers = hip.Walkers(self.ctx, x, states)
            self.walkers.start(stream=self.stream, delta=self.delta)
            if self.single:
                self.sm = hip.SingleMoves(self.walkers)
                self.sm.refresh(stream=self.stream)
        self.nblk = 0
        self.nsweep = 0
        self.blocks = []
        self.evals = 0
        self.event = torch.cuda.Event()

    def queue_block(self):
        import torch
        with torch.cuda.stream(self.stream):
            self.walkers.blk.zero_()
            self.walkers.sweep(self.spec.steps_per_block, self.delta, mage=self.spec.mage,
                               stream=self.stream)
            self.event.record(self.stream)

    def finish_block(self):
        """Wait for the queued block; fold its mean into the statistics.
        Returns True when the job has reached its target."""
        import torch
        self.event.synchronize()
        with torch.cuda.stream(self.stream):
            e = self.walkers.blk[:, self.jetot].sum().item()
        self.evals += self.spec.nw * self.spec.steps_per_block
        mean = e / (self.spec.nw * self.spec.steps_per_block)
        self.nblk += 1
        if self.nblk > self.spec.equil_blocks:
            self.blocks.append(mean)
        n = len(self.blocks)
        if n >= self.spec.max_blocks:
            return True
        return n >= self.spec.min_blocks and block_error(self.blocks) <= self.spec.target

    def result(self, t_done):
        st = self.walkers.status()
        if np.any(st & 2):
            raise RuntimeError(f"job {self.spec.job_id}: singular orbital matrix")
        err = block_error(self.blocks)
        held = self.ctx.device_bytes()["dev_bytes"] + hip.Walkers.state_bytes(
            self.spec.nw, self.deck.nptot, self.ctx.info["n_props"])
        if self.single:
            held += self.sm.ainv.numel() * self.sm.ainv.element_size() + 2 * 8 * self.spec.nw
        return JobResult(self.spec.job_id, float(np.mean(self.blocks)) if self.blocks else math.nan,
                         err, list(self.blocks), self.spec.arrival, self.t_admit, t_done,
                         self.evals, err <= self.spec.target, self.need, held)

    def close(self):
        self.stream.synchronize()
        self.sm = None
        self.walkers = None
        self.ctx.close()


def serve(specs, mem_budget=64 << 30, clock=None, max_active=64, poll_s=0.0005, moves="all"):
    """Run the jobs to their error targets with continuous batching.
    Returns (results by job id, summary dict).  The latency percentiles are
    over the jobs that reached their target; `capped` counts the jobs that
    stopped at max_blocks without it.

    moves="all": each job runs the reference's all-electron VMC step on a
    stream of its own.  moves="single" (jobs without backflow): every
    electron of every running job is moved once per sweep by packed
    single-electron launches (up to 16 jobs of any N per launch), then each
    job's walkers are evaluated for their local energies; a block is
    steps_per_block sweeps, and jobs are admitted between sweeps."""
    if moves == "single":
        return _serve_single(specs, mem_budget, clock, max_active)
    if moves != "all":
        raise ValueError(f"unknown moves {moves!r}")
    import torch
    clock = clock or time.perf_counter
    pending = sorted(specs, key=lambda s: (s.arrival, s.job_id))
    workdir = tempfile.mkdtemp(prefix="degas_serve_")
    active, results = [], {}
    used = 0
    head = None   # the next job, its context created, waiting for memory
    t0 = clock()
    blocks_run = 0
    while pending or active:
        now = clock() - t0
        # admission in arrival order while the job's memory fits
        while pending and pending[0].arrival <= now and len(active) < max_active:
            if head is None:
                head = _Running(pending[0], workdir)
            if active and used + head.need > mem_budget:
                break
            pending.pop(0)
            r, head = head, None
            r.start(clock() - t0)
            used += r.need
            r.queue_block()
            active.append(r)
        if not active:
            time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
            continue
        progressed = False
        still = []
        for r in active:
            if not r.event.query():
                still.append(r)
                continue
            progressed = True
            blocks_run += 1
            if r.finish_block():
                res = r.result(clock() - t0)
                results[res.job_id] = res
                used -= r.need
                r.close()
            else:
                r.queue_block()
                still.append(r)
        active = still
        if not progressed:
            time.sleep(poll_s)
    torch.cuda.synchronize()
    return results, _summary(results, clock() - t0, blocks_run)


def _serve_single(specs, mem_budget, clock, max_active):
    import torch
    clock = clock or time.perf_counter
    pending = sorted(specs, key=lambda s: (s.arrival, s.job_id))
    workdir = tempfile.mkdtemp(prefix="degas_serve_")
    active, results = [], {}
    used = 0
    head = None
    t0 = clock()
    blocks_run = 0
    while pending or active:
        now = clock() - t0
        while pending and pending[0].arrival <= now and len(active) < max_active:
            if head is None:
                head = _Running(pending[0], workdir, single=True)
            if active and used + head.need > mem_budget:
                break
            pending.pop(0)
            r, head = head, None
            r.start(clock() - t0)
            r.walkers.blk.zero_()
            used += r.need
            active.append(r)
        if not active:
            time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
            continue
        # one sweep of every running job, 16 jobs per packed call
        for g in range(0, len(active), 16):
            hip.sem_moves_packed([(r.sm, 0, r.deck.nptot, r.spec.sigma) for r in active[g:g + 16]])
        for r in active:
            r.walkers.start()
            r.walkers.blk += r.walkers.props
            r.nsweep += 1
        still = []
        for r in active:
            if r.nsweep < r.spec.steps_per_block:
                still.append(r)
                continue
            r.nsweep = 0
            blocks_run += 1
            r.event.record()
            if r.finish_block():
                res = r.result(clock() - t0)
                results[res.job_id] = res
                used -= r.need
                r.close()
            else:
                r.walkers.blk.zero_()
                r.sm.refresh()
                still.append(r)
        active = still
    torch.cuda.synchronize()
    return results, _summary(results, clock() - t0, blocks_run)


def _summary(results, wall, blocks_run):
    lat = np.array([r.latency for r in results.values() if r.converged])
    evals = sum(r.evals for r in results.values())
    summary = {
        "jobs": len(results),
        "capped": int(sum(not r.converged for r in results.values())),
        "wall_s": wall,
        "jobs_per_s": len(results) / wall if wall > 0 else 0.0,
        "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
        "evals_per_s": evals / wall if wall > 0 else 0.0,
        "blocks": blocks_run,
        "converged": int(sum(r.converged for r in results.values())),
    }
    return summary


def poisson_jobs(n, rate, seed, nw=256, target=2e-4, sizes=((13, 13), (29, 29), (37, 37), (61, 61),
                                                           (37, 0), (49, 0), (61, 0)),
                 rs_range=(1.0, 20.0), max_blocks=400, steps_per_block=20, backflow="yk"):
    """n job specs with Poisson arrivals at `rate` per second, N drawn from
    `sizes` (paramagnetic and polarized), rs uniform in rs_range, and the
    wavefunction parameters perturbed by 10% from a per-job seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n):
        t += rng.exponential(1.0 / rate)
        nu, nd = sizes[rng.integers(len(sizes))]
        rs = float(rng.uniform(*rs_range))
        out.append(JobSpec(job_id=j, nup=int(nu), ndown=int(nd), rs=rs,
                           param_seed=int(rng.integers(1, 1 << 30)), nw=nw, target=target,
                           arrival=t, max_blocks=max_blocks, steps_per_block=steps_per_block,
                           backflow=backflow))
    return out
