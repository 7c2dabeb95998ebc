"""Host logic of the job service (2degas_amd/serve.py): the reference's
block error, the walker-state bytes admission counts and the arrival
generator."""
import math

import numpy as np

from conftest import pkg


def test_block_error_matches_reference_formula():
    """averages(3): err = sqrt(|cml2/n - (cml/n)^2| / (n - 1))
    (reptation.f:4336-4346), written out for four block means."""
    sv = pkg("serve")
    b = [1.0, 1.1, 0.9, 1.05]
    n = 4
    cml, cml2 = sum(b), sum(x * x for x in b)
    ref = math.sqrt(abs(cml2 / n - (cml / n) ** 2) / (n - 1))
    assert abs(sv.block_error(b) - ref) < 1e-15
    # the standard error of the mean of independent blocks
    assert abs(sv.block_error(b) - np.std(b, ddof=1) / math.sqrt(n)) < 1e-12
    assert math.isinf(sv.block_error([1.0]))


def test_walker_state_bytes_are_the_tensors_walkers_allocates():
    """hip.Walkers.state_bytes: x and g (nw x N x 2 doubles each), h (nw x 2),
    props and block sums (nw x n_props each), grad2_ph, age and the RNG
    state (one 8-byte word each per walker)."""
    hip = pkg("hip")
    nw, N, npr = 256, 122, 40
    want = nw * 8 * (2 * N * 2 + 2 + npr + 1 + 1 + 1 + npr)
    assert hip.Walkers.state_bytes(nw, N, npr) == want
    assert hip.Walkers.state_bytes(nw, N, npr, block_sums=False) == want - nw * 8 * npr


def test_poisson_arrivals_and_job_mix():
    sv = pkg("serve")
    jobs = sv.poisson_jobs(2000, 20.0, seed=5)
    t = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.1 / 20.0
    rs = np.array([j.rs for j in jobs])
    assert rs.min() >= 1.0 and rs.max() <= 20.0
    sizes = {(j.nup, j.ndown) for j in jobs}
    assert (61, 61) in sizes and (37, 0) in sizes
    assert len({j.param_seed for j in jobs}) > 1900
