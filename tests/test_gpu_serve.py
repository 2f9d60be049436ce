"""Served mode: a request stream through continuous batching gives, for every
request, the bits the same system gives alone.

Requests arrive (seeded Poisson stream, fast enough that they overlap) at a
pool with room for about three of them; each is admitted when its footprint
fits, stepped with whatever else is resident, harvested and removed.  The
digest of each result (every particle's disp, CurrPos, vel, acc, force,
intact bonds, wall planes and reaction) must equal the digest of the same
request run alone through the plain interface.  Tolerance: 0 ulp.
"""
import pytest

from helpers import small_jobs
from peridem_amd import serve

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    systems = small_jobs(10, seed=41)
    budgets = [35, 20, 48, 27, 60, 18, 41, 33, 25, 52]
    arrivals = serve.poisson_arrivals(len(systems), rate=400.0, seed=3)
    alone_t, alone_d = serve.run_sequential(systems, budgets, digests=True)
    cap = serve.pool_capacity(systems[:3], slack=1.2)
    srv = serve.Server(chunk=16, **cap)
    try:
        done = srv.serve(list(zip(arrivals, systems, budgets)))
    finally:
        srv.close()
    assert sorted(d.rid for d in done) == list(range(len(systems)))
    assert srv.peak_live > 1, "requests never overlapped on the device"
    for d in done:
        assert d.digest == alone_d[d.rid], f"request {d.rid} differs from its solo run"
        assert d.latency >= 0 and d.finished >= d.admitted >= 0


def test_batch_digests_equal_alone():
    systems = small_jobs(6, seed=43)
    _, _, batch_d, _ = serve.run_batch(systems, 30, digests=True)
    _, alone_d = serve.run_sequential(systems, 30, digests=True)
    assert batch_d == alone_d
