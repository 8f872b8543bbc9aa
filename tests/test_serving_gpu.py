"""Serving engine (pcf_submit / pcf_poll / pcf_wait): a mixed stream of
requests over resident catalogs gives, for every request, the result of the
same histogram computed alone (counts bit-exact, weights to 1e-12 relative),
with concurrent lanes, continuous batching and HBM admission."""
import numpy as np
import pytest

import desi_amd as da
from desi_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture()
def served():
    e = da.Engine(0)
    yield e
    e.close()


def _catalogs(e):
    cats = {
        "u4k": synth.uniform(4000, 250.0, seed=201, weights="uniform"),
        "u9k": synth.uniform(9000, 250.0, seed=202),
        "ns3k": synth.neyman_scott(100, 30, 5.0, 250.0, seed=203, weights="uniform"),
        "u1k": synth.uniform(1200, 250.0, seed=204),
        "big": synth.uniform(60000, 1000.0, seed=205),
        "big2": synth.uniform(40000, 1000.0, seed=206, weights="uniform"),
    }
    return {k: e.upload(v) for k, v in cats.items()}


def _mix(ids):
    J = da.make_job
    jobs = []
    for bins in (20, 30):
        for bpc in (False, True):
            jobs.append(J(da.PCF_2ISO, da.PCF_XX, bins=bins, dmax=60.0, cat_x=ids["u4k"], bpc=bpc,
                          box=250.0))
            jobs.append(J(da.PCF_2ISO, da.PCF_XY, bins=bins, dmax=60.0, cat_x=ids["u4k"],
                          cat_y=ids["u9k"], bpc=bpc, box=250.0))
            jobs.append(J(da.PCF_2ISO, da.PCF_XY, bins=bins, dmax=60.0, cat_x=ids["ns3k"],
                          cat_y=ids["u9k"], bpc=bpc, box=250.0))
            jobs.append(J(da.PCF_2ANI, da.PCF_XY, bins=bins, dmax=80.0, cat_x=ids["ns3k"],
                          cat_y=ids["u4k"], bpc=bpc, box=250.0))
            jobs.append(J(da.PCF_2ANI, da.PCF_XX, bins=bins, dmax=80.0, cat_x=ids["u9k"], bpc=bpc,
                          box=250.0))
    for bpc in (False, True):
        jobs.append(J(da.PCF_3ISO, da.PCF_XXX, bins=20, dmax=40.0, cat_x=ids["u1k"], bpc=bpc,
                      box=250.0))
        jobs.append(J(da.PCF_3ISO, da.PCF_XXY, bins=20, dmax=40.0, cat_x=ids["u1k"],
                      cat_y=ids["ns3k"], bpc=bpc, box=250.0))
        jobs.append(J(da.PCF_3ANI, da.PCF_XXX, bins=8, dmax=30.0, cat_x=ids["u1k"], bpc=bpc,
                      box=250.0))
    jobs.append(J(da.PCF_2ISO, da.PCF_XY, bins=30, dmax=80.0, cat_x=ids["big"], cat_y=ids["big2"]))
    jobs.append(J(da.PCF_2ISO, da.PCF_XX, bins=30, dmax=80.0, cat_x=ids["big2"]))
    jobs.append(J(da.PCF_2ISO, da.PCF_XY, bins=30, dmax=60.0, cat_x=ids["u4k"], cat_y=ids["u9k"],
                  precision=da.PCF_PREC_FP64))
    return jobs


def _check_mix(e, jobs, results):
    for k, (j, (c, w, st)) in enumerate(zip(jobs, results)):
        c1, w1 = e.histogram(j)
        assert np.array_equal(c, c1), f"request {k}: counts differ in the mix"
        assert np.all(np.abs(w - w1) <= 1e-12 * np.abs(w1)), f"request {k}: weights"
        assert st["latency_ms"] >= st["queue_ms"] >= 0


@pytest.mark.parametrize("lanes,max_batch", [(1, 1), (4, 1), (4, 6)])
def test_mix_equals_alone(served, lanes, max_batch):
    e = served
    ids = _catalogs(e)
    e.set_hbm_budget(0, lanes)
    e.set_max_batch(max_batch)
    jobs = _mix(ids)
    print(f"case serving mix lanes={lanes} max_batch={max_batch} requests={len(jobs)}",
          flush=True)
    tickets = [e.submit(j) for j in jobs]
    results = [e.wait(t, j.type, j.bins) for t, j in zip(tickets, jobs)]
    if max_batch > 1:
        assert max(r[2]["batched_jobs"] for r in results) > 1
    _check_mix(e, jobs, results)


def test_heterogeneous_batches_in_the_mix(served):
    """Pair requests of one type and box with different bins and dmax queue
    behind a long job on one lane and are served as one launch: each result
    equals the request alone."""
    e = served
    ids = _catalogs(e)
    e.set_hbm_budget(0, 1)
    e.set_max_batch(8)
    J = da.make_job
    first = J(da.PCF_2ISO, da.PCF_XY, bins=30, dmax=80.0, cat_x=ids["big"], cat_y=ids["big2"])
    jobs = [J(da.PCF_2ISO, da.PCF_XX, bins=b, dmax=dm, cat_x=ids[x], bpc=True, box=250.0)
            for b, dm, x in ((20, 45.0, "u4k"), (30, 60.0, "ns3k"), (40, 97.5, "u9k"),
                             (30, 120.0, "u1k"))]
    jobs += [J(da.PCF_2ISO, da.PCF_XY, bins=b, dmax=dm, cat_x=ids["ns3k"], cat_y=ids["u4k"],
               bpc=True, box=250.0) for b, dm in ((20, 70.0), (40, 52.0))]
    t0 = e.submit(first)
    tickets = [e.submit(j) for j in jobs]
    e.wait(t0, first.type, first.bins)
    results = [e.wait(t, j.type, j.bins) for t, j in zip(tickets, jobs)]
    assert max(r[2]["batched_jobs"] for r in results) > 1
    _check_mix(e, jobs, results)


def test_tiny_budget_serialises(served):
    e = served
    ids = _catalogs(e)
    e.set_hbm_budget(1, 4)  # every job exceeds it: one job at a time
    jobs = _mix(ids)[:12]
    tickets = [e.submit(j) for j in jobs]
    res = [e.wait(t, j.type, j.bins) for t, j in zip(tickets, jobs)]
    spans = sorted((r[2]["queue_ms"], r[2]["latency_ms"]) for r in res)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 0.5, "jobs overlapped under a budget no job fits"
    _check_mix(e, jobs, res)


def test_poll_wait_errors_and_release(served):
    e = served
    ids = _catalogs(e)
    j = da.make_job(da.PCF_2ISO, da.PCF_XY, bins=20, dmax=60.0, cat_x=ids["u4k"], cat_y=ids["u9k"])
    t = e.submit(j)
    c, w, st = e.wait(t, j.type, j.bins)
    with pytest.raises(da.PcfError):
        e.poll(t)  # collected
    bad = da.make_job(da.PCF_2ISO, da.PCF_XX, bins=20, dmax=260.0, cat_x=ids["u4k"], bpc=True,
                      box=250.0)  # dmax >= box
    tb = e.submit(bad)
    with pytest.raises(da.PcfError):
        e.wait(tb, bad.type, bad.bins)
    # releasing a catalog waits for the queued jobs that use it
    ts = [e.submit(j) for _ in range(6)]
    e.release(ids["u9k"])
    for tt in ts:
        assert e.poll(tt)
        c2, _, _ = e.wait(tt, j.type, j.bins)
        assert np.array_equal(c2, c)


def test_refused_job_does_not_fail_its_batch(served):
    """A job refused on its own arguments (a periodic job whose catalog lies
    outside the box) queued with compatible jobs: they still get their own
    results, and it gets its error."""
    e = served
    ids = _catalogs(e)
    e.set_hbm_budget(0, 1)  # one lane: jobs queue behind the first and batch
    e.set_max_batch(8)
    J = da.make_job
    first = J(da.PCF_2ISO, da.PCF_XY, bins=30, dmax=80.0, cat_x=ids["big"], cat_y=ids["big2"])
    kw = dict(bins=20, dmax=60.0, bpc=True, box=250.0)
    good1 = J(da.PCF_2ISO, da.PCF_XX, cat_x=ids["u4k"], **kw)
    bad = J(da.PCF_2ISO, da.PCF_XX, cat_x=ids["big"], **kw)  # points up to 1000
    good2 = J(da.PCF_2ISO, da.PCF_XY, cat_x=ids["ns3k"], cat_y=ids["u9k"], **kw)
    t0 = e.submit(first)
    tg1, tb, tg2 = e.submit(good1), e.submit(bad), e.submit(good2)
    e.wait(t0, first.type, first.bins)
    for t, j in ((tg1, good1), (tg2, good2)):
        c, w, _ = e.wait(t, j.type, j.bins)
        c1, w1 = e.histogram(j)
        assert np.array_equal(c, c1)
        assert np.all(np.abs(w - w1) <= 1e-12 * np.abs(w1))
    with pytest.raises(da.PcfError) as err:
        e.wait(tb, bad.type, bad.bins)
    assert err.value.rc == da.PCF_ERR_RANGE


def test_budget_change_keeps_finished_tickets(served):
    e = served
    ids = _catalogs(e)
    J = da.make_job
    jobs = [J(da.PCF_2ISO, da.PCF_XY, bins=20, dmax=60.0, cat_x=ids["u4k"], cat_y=ids["u9k"]),
            J(da.PCF_2ANI, da.PCF_XX, bins=20, dmax=60.0, cat_x=ids["ns3k"])]
    tickets = [e.submit(j) for j in jobs]
    e.set_hbm_budget(0, 2)  # restarts the engine after the queue drains
    for t, j in zip(tickets, jobs):
        c, _, _ = e.wait(t, j.type, j.bins)
        c1, _ = e.histogram(j)
        assert np.array_equal(c, c1)
    t = e.submit(jobs[0])  # ticket ids keep counting on the new engine
    assert t > max(tickets)
    e.wait(t, jobs[0].type, jobs[0].bins)
