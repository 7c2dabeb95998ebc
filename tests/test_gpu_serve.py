"""The job service on the device: every job's block means are the same run
alone and run in a mix of jobs that arrive while it runs."""
import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def _specs(sv):
    # fixed block counts (target 0, max_blocks) so every run is comparable
    common = dict(nw=64, steps_per_block=4, target=0.0, min_blocks=4, max_blocks=5, shells=10)
    return [sv.JobSpec(0, 13, 13, 5.0, param_seed=11, arrival=0.0, **common),
            sv.JobSpec(1, 29, 29, 2.0, param_seed=12, arrival=0.05, **common),
            sv.JobSpec(2, 37, 0, 8.0, param_seed=13, arrival=0.1, **common)]


def test_jobs_give_the_same_results_alone_and_in_the_mix():
    print("\n[gpu] serve alone vs mix", flush=True)
    sv = pkg("serve")
    alone = {}
    for s in _specs(sv):
        s.arrival = 0.0
        res, _ = sv.serve([s])
        alone[s.job_id] = res[s.job_id].blocks
    mixed, summary = sv.serve(_specs(sv))
    print(f"   summary {summary}", flush=True)
    assert summary["jobs"] == 3
    for j, blocks in alone.items():
        assert len(blocks) == 5
        assert np.all(np.isfinite(blocks))
        assert mixed[j].blocks == blocks, f"job {j} differs in the mix"
    # every job stopped at max_blocks: none reached its target, so there are
    # no time-to-target latencies
    assert summary["capped"] == 3 and summary["converged"] == 0
    assert summary["p50_latency_s"] is None


def test_admission_counts_what_each_job_allocates():
    """Each job's admission bytes (qmc_ctx_info's sizes for its walker count
    plus its walker tensors) equal what its context and walkers held when it
    retired.  With a budget below any one job, jobs run one at a time in
    arrival order; with a reachable target every job converges and the
    latency percentiles are over all of them."""
    print("\n[gpu] serve admission by allocation", flush=True)
    sv = pkg("serve")
    specs = _specs(sv)
    for s in specs:
        s.target, s.min_blocks, s.max_blocks = 10.0, 2, 5
    res, summary = sv.serve(specs, mem_budget=1)
    print(f"   summary {summary}", flush=True)
    order = sorted(res.values(), key=lambda r: r.admitted)
    for r in order:
        print(f"   job {r.job_id}: predicted {r.predicted_bytes} held {r.device_bytes}", flush=True)
        assert r.predicted_bytes == r.device_bytes > 0
        assert r.converged and len(r.blocks) == 2
    assert [r.job_id for r in order] == [0, 1, 2]
    for a, b in zip(order, order[1:]):
        assert b.admitted >= a.done
    assert summary["capped"] == 0 and summary["converged"] == 3
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0


def _single_specs(sv):
    common = dict(nw=64, steps_per_block=4, target=0.0, min_blocks=4, max_blocks=5, shells=10,
                  backflow="zero", u3=True)
    return [sv.JobSpec(0, 13, 13, 5.0, param_seed=21, arrival=0.0, **common),
            sv.JobSpec(1, 29, 29, 2.0, param_seed=22, arrival=0.02, **common),
            sv.JobSpec(2, 37, 0, 8.0, param_seed=23, arrival=0.04, sigma=0.4, **common)]


def test_single_electron_service_alone_and_in_the_mix():
    """moves="single": jobs without backflow share packed single-electron
    launches; each job's block means are the same alone and in the mix, and
    its admission bytes (with the single-electron state) equal what it held
    when it retired."""
    print("\n[gpu] serve single-electron moves, alone vs mix", flush=True)
    sv = pkg("serve")
    alone = {}
    for s in _single_specs(sv):
        s.arrival = 0.0
        res, _ = sv.serve([s], moves="single")
        alone[s.job_id] = res[s.job_id].blocks
    mixed, summary = sv.serve(_single_specs(sv), moves="single")
    print(f"   summary {summary}", flush=True)
    assert summary["jobs"] == 3
    for j, blocks in alone.items():
        assert len(blocks) == 5
        assert np.all(np.isfinite(blocks))
        assert mixed[j].blocks == blocks, f"job {j} differs in the mix"
        assert mixed[j].predicted_bytes == mixed[j].device_bytes > 0
    with pytest.raises(ValueError):
        s = _single_specs(sv)[0]
        s.backflow = "yk"
        sv.serve([s], moves="single")
