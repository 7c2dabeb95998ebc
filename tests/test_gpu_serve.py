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
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
