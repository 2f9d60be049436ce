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


def test_c_request_loop_results_equal_alone():
    """The library's own request loop (pd_serve_submit / pd_serve_poll /
    pd_serve_release, for C and C++ hosts): every request's results equal its
    solo run bit for bit; a request larger than the empty pool comes back
    with PD_E_NOMEM, one whose contact lists would need more than 255 entries
    with PD_E_OVERFLOW, and neither disturbs the others."""
    import numpy as np

    from peridem_amd import abi, gen
    systems = small_jobs(8, seed=47)
    budgets = [30, 18, 44, 25, 52, 16, 38, 29]
    _, alone_d = serve.run_sequential(systems, budgets, digests=True)
    h = 1e-4
    big = gen.make_particle(gen.lattice_sphere(5.0), h, np.zeros(3), np.random.default_rng(41))
    stopped = gen.System(particles=[big], dt=1e-8, contact_rad=3.5 * h,
                         normal_stiffness=gen.normal_stiffness(72e9, 0.25, 3.5 * h), self_contact=1,
                         enable_fracture=1)
    huge = gen.two_sphere_impact(R=1.2e-3, h=1e-4, v=10.0, seed=2)
    cap = serve.pool_capacity(systems[:3] + [stopped], slack=1.2)
    pool = abi.Pool(cap["max_nodes"], cap["max_bond_entries"], cap["max_particles"], cap["max_jobs"])
    srv = abi.CServer(pool, chunk=12)
    try:
        rid = {}
        for k, (s, n) in enumerate(zip(systems, budgets)):
            rid[srv.submit(s, n)] = k
            if k == 2:
                r_stop = srv.submit(stopped, 20)
                r_huge = srv.submit(huge, 10)
        got, status, peak = {}, {}, 0
        while len(status) < len(systems) + 2:
            for e in srv.poll():
                status[e["request"]] = e["status"]
                assert e["t_done"] >= e["t_admit"] >= e["t_submit"] >= 0
                if e["job"] != 0xFFFFFFFF:
                    if e["status"] == 0:
                        k = rid[e["request"]]
                        got[k] = serve.digest_state(lambda p, j=e["job"]: pool.state(j, p),
                                                    lambda p, j=e["job"]: pool.connectivity(j, p),
                                                    lambda j=e["job"]: pool.wall(j), len(systems[k].particles))
                    srv.release(e["job"])
            peak = max(peak, srv.pending()[1])
        assert status[r_huge] == -6 and status[r_stop] == -5
        assert peak > 1, "requests never overlapped on the device"
        for k in range(len(systems)):
            assert got[k] == alone_d[k], f"request {k} differs from its solo run"
    finally:
        srv.close()
        pool.close()
