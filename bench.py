#!/usr/bin/env python3
"""Benchmark of the grid-node pair counting hot path on one MI355X.

Workload (the north-star 2PCF shape): 2iso, data and random catalogs of
N = 1,048,576 points each, uniform FP32 in a box of side 1000 (seeded), rmax
80, 30 bins, no periodic boundaries.  One step = the hot path over one batch,
in the batched form: one pcf_histogram_batch call (a replayed HIP graph) that
re-runs the device grid build of both catalogs (counting sort into cell
order) and then the DD, RR and DR histograms in one persistent launch of the
FP32 + FP64-edge-recheck kernel.  Inputs are resident in HBM before the timed
region (catalog upload is not timed).

Every leg checks its mixed-precision histograms bin by bin against the plain
FP64 GPU path (counts equal, weighted sums within 1e-12 relative) and reports
the largest relative error of the weighted sums per histogram.

value = useful pairs per second (sum of the three unweighted histograms divided
by the wall time of the step), the reference's "pairs/s".  The pair kernel is
the autotuned choice: for this unweighted 2iso workload, normally the Gram form
(d2 from the matrix cores) over the direct form (d2 on the vector units).
Also reported: the same step on the plain FP64 path, the roofline of the pair
kernel (8 FP32 flops per candidate-pair evaluation against the 157.3 TFLOP/s
FP32 vector peak), and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import desi_amd as da  # noqa: E402
from desi_amd import synth  # noqa: E402

PEAK_FP32 = 157.3e12  # MI355X FP32 peak, FLOP/s (packed vector and dense matrix alike)
FLOPS_PER_PAIR = 8    # 3 sub + 3 mul + 2 add, or 4 chained FMAs in Gram form (FMA = 2 flops)
FLOPS_PER_TRIPLET = 16  # two squared distances per third-point evaluation
N = 1 << 20
BOX = 1000.0
RMAX = 80.0
BINS = 30


# rocprofv3 summaries of the NS pair kernel: the Gram form (matrix cores) and
# the direct form (vector units); the tuned choice's summary gives `traffic`.
# Counter summary of the tuned pair kernel by form (the Gram form is the one the autotuner
# picks on this workload; another pick reports traffic null).
PAIR_PROFILE = {2: os.path.join(ROOT, "profiles", "k_pairs_gram_2iso_N1M_box1000_rmax80.md")}
# HBM traffic of the triplet and batch legs' counting kernels (average bytes per
# dispatch of the timed step), from their committed counter summaries.
ISO_PROFILE = os.path.join(ROOT, "profiles", "k_triplets_3iso_bpc_N32K_box250_rmax80.md")
ANI_PROFILE = os.path.join(ROOT, "profiles", "k_trip_ani_lds_3ani_N64K_box500_rmax60.md")
BATCH_PROFILE = os.path.join(ROOT, "profiles", "k_pairs_mixed_batch_2ani_64randoms.md")


def profiled_traffic(path, kernel):
    """HBM bytes per dispatch of `kernel` from a committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the reading of a plain
    copy of known size), or None when the summary is absent."""
    if path is None:
        return None
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return float(line.split(":")[-1])
    except OSError:
        pass
    return None


WREL = 1e-12  # relative tolerance of weighted sums against the plain FP64 path


def compare(name, mixed, plain):
    """Per-bin parity of one histogram: counts equal, weights' largest
    relative error.  Raises on a count mismatch or an error above WREL."""
    (c, w), (c64, w64) = mixed, plain
    if not np.array_equal(c, c64):
        bad = np.nonzero(c != c64)[0]
        raise SystemExit(f"{name}: mixed and plain FP64 counts differ in {len(bad)} bins "
                         f"(first {bad[:5].tolist()})")
    den = np.maximum(np.abs(w64), 1e-300)
    err = float(np.max(np.abs(w - w64) / den)) if len(w) else 0.0
    if err > WREL:
        raise SystemExit(f"{name}: weighted sums off by {err:.3e} relative")
    return {"bins": int(len(c)), "counts_equal": True, "max_rel_err_w": err,
            "total": int(c.sum())}


def ns_jobs(cd, cr, precision=da.PCF_PREC_MIXED, rebuild=True):
    """DD, RR and DR of the NS step; rebuild_grid makes every launch re-run the
    device grid build of both catalogs."""
    return [da.make_job(da.PCF_2ISO, da.PCF_XX, bins=BINS, dmax=RMAX, cat_x=cd,
                        precision=precision, rebuild_grid=rebuild),
            da.make_job(da.PCF_2ISO, da.PCF_XX, bins=BINS, dmax=RMAX, cat_x=cr,
                        precision=precision, rebuild_grid=rebuild),
            da.make_job(da.PCF_2ISO, da.PCF_XY, bins=BINS, dmax=RMAX, cat_x=cd, cat_y=cr,
                        precision=precision, rebuild_grid=rebuild)]


def run_batched_step(e, jobs):
    res = e.histogram_batch(jobs)
    st = e.stats()
    return res, st


def run_step(e, cd, cr, precision, rebuild=True):
    tot_pairs = 0
    evals = 0
    kms = 0.0
    rech = 0
    # DD re-grids D, RR re-grids R; DR uses the two grids just built (the
    # three share one geometry).
    for hist, cy, rb in ((da.PCF_XX, -1, rebuild), (da.PCF_XX, None, rebuild),
                         (da.PCF_XY, cr, False)):
        cx = cr if cy is None else cd
        job = da.make_job(da.PCF_2ISO, hist, bins=BINS, dmax=RMAX, cat_x=cx,
                          cat_y=-1 if cy is None else cy, precision=precision,
                          rebuild_grid=rb)
        c, _ = e.histogram(job)
        st = e.stats()
        tot_pairs += int(c.sum())
        evals += st["evaluations"]
        kms += st["kernel_ms"]
        rech += st["rechecks"]
    return tot_pairs, evals, kms, rech


def run_step_hists(e, cd, cr, precision):
    """The NS histograms as separate calls: [(counts, weights)] for DD, RR, DR."""
    out = []
    for j in ns_jobs(cd, cr, precision, rebuild=False):
        out.append(e.histogram(j))
    return out


TRIP_NAMES = ("DDD", "DDR", "DRR", "RRR")


def run_triplet_step(e, cd, cr, precision, rebuild=True):
    """Config 2: 3iso with periodic boundaries, DDD + DDR + DRR + RRR."""
    tot = 0
    evals = 0
    kms = 0.0
    hists = []
    for hist, x, y in ((da.PCF_XXX, cd, -1), (da.PCF_XXY, cd, cr), (da.PCF_XXY, cr, cd),
                       (da.PCF_XXX, cr, -1)):
        job = da.make_job(da.PCF_3ISO, hist, bins=30, dmax=80.0, cat_x=x, cat_y=y, bpc=True,
                          box=250.0, precision=precision, rebuild_grid=rebuild)
        c, w = e.histogram(job)
        hists.append((c, w))
        st = e.stats()
        tot += int(c.sum())
        evals += st["evaluations"]
        kms += st["kernel_ms"]
    return tot, evals, kms, hists


def triplet_ani_leg(e, steps):
    """3ani periodic: N=65,536 Neyman-Scott (2048 parents x 32, sigma 4) and
    65,536 uniform, box 500, rmax 60, 15 bins, weights U[0.5,1.5]: DDD + DDR +
    DRR + RRR, mixed path against the plain FP64 path, each warmed by one
    untimed step and then timed over `steps` steps."""
    d = synth.neyman_scott(2048, 32, 4.0, 500.0, seed=41, weights="uniform")
    r = synth.uniform(65536, 500.0, seed=42, weights="uniform")
    cd, cr = e.upload(d), e.upload(r)
    out = {}
    hists = {}

    def step(prec):
        tot = evals = 0
        kms = 0.0
        hs = []
        for hist, x, y in ((da.PCF_XXX, cd, -1), (da.PCF_XXY, cd, cr), (da.PCF_XXY, cr, cd),
                           (da.PCF_XXX, cr, -1)):
            job = da.make_job(da.PCF_3ANI, hist, bins=15, dmax=60.0, cat_x=x, cat_y=y, bpc=True,
                              box=500.0, precision=prec)
            c, w = e.histogram(job)
            hs.append((c, w))
            st = e.stats()
            tot += int(c.sum())
            evals += st["evaluations"]
            kms += st["kernel_ms"]
        return tot, evals, kms, hs

    for prec, name in ((da.PCF_PREC_MIXED, "mixed"), (da.PCF_PREC_FP64, "fp64_plain")):
        step(prec)
        tot = evals = 0
        kms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            t, ev, k, hists[name] = step(prec)
            tot += t
            evals += ev
            kms += k
        t1 = time.perf_counter()
        out[name] = {"ms": (t1 - t0) * 1e3 / steps, "kernel_ms": kms / steps,
                     "triplets": tot // steps, "evaluations": evals // steps}
    e.release(cd)
    e.release(cr)
    parity = {n: compare(f"3ani {n}", hm, hp)
              for n, hm, hp in zip(TRIP_NAMES, hists["mixed"], hists["fp64_plain"])}
    m = out["mixed"]
    ach = m["evaluations"] * (FLOPS_PER_TRIPLET + 2) / (m["kernel_ms"] * 1e-3)
    return {"workload": "3ani periodic box 500, N=65536 Neyman-Scott (2048x32, sigma 4) + 65536 "
                        "uniform, weights U[0.5,1.5], rmax 60, 15 bins, DDD+DDR+DRR+RRR",
            "value": m["triplets"] / (m["ms"] * 1e-3), "unit": "triplets/s",
            "ms_per_step": m["ms"], "steps": steps, "warmup": 1,
            "kernel_ms_per_step": m["kernel_ms"],
            "evaluations_per_step": m["evaluations"],
            "fp64_plain_ms_per_step": out["fp64_plain"]["ms"],
            "fp64_plain_kernel_ms_per_step": out["fp64_plain"]["kernel_ms"],
            "speedup_vs_fp64_plain": out["fp64_plain"]["ms"] / m["ms"],
            "parity_vs_fp64": parity,
            "roofline": {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                         "traffic": profiled_traffic(ANI_PROFILE, "k_trip_ani_lds 3ani")}}


def triplet_leg(e, steps):
    d = synth.neyman_scott(1024, 32, 5.0, 250.0, seed=11, weights="uniform")
    r = synth.uniform(32768, 250.0, seed=12, weights="uniform")
    cd, cr = e.upload(d), e.upload(r)
    run_triplet_step(e, cd, cr, da.PCF_PREC_MIXED)
    t0 = time.perf_counter()
    tot = evals = 0
    kms = 0.0
    for _ in range(steps):
        t, ev, k, hm = run_triplet_step(e, cd, cr, da.PCF_PREC_MIXED)
        tot += t
        evals += ev
        kms += k
    t1 = time.perf_counter()
    run_triplet_step(e, cd, cr, da.PCF_PREC_FP64)  # warm-up, like the mixed path's
    t1b = time.perf_counter()
    t64, _, k64, hp = run_triplet_step(e, cd, cr, da.PCF_PREC_FP64)
    t2 = time.perf_counter()
    parity = {n: compare(f"3iso {n}", a, b) for n, a, b in zip(TRIP_NAMES, hm, hp)}
    e.release(cd)
    e.release(cr)
    ms = (t1 - t0) * 1e3 / steps
    ach = evals * FLOPS_PER_TRIPLET / (kms * 1e-3)
    return {
        "workload": "3iso periodic box 250, N=32768 Neyman-Scott (1024x32, sigma 5) + 32768 "
                    "uniform, weights U[0.5,1.5], rmax 80, 30 bins, DDD+DDR+DRR+RRR",
        "value": tot / ((t1 - t0) / steps), "unit": "triplets/s",
        "ms_per_step": ms, "steps": steps,
        "kernel_ms_per_step": kms / steps,
        "evaluations_per_step": evals // steps,
        "fp64_plain_ms_per_step": (t2 - t1b) * 1e3,
        "fp64_plain_kernel_ms_per_step": k64,
        "speedup_vs_fp64_plain": (t2 - t1b) * 1e3 / ms,
        "parity_vs_fp64": parity,
        # the 2-3 side of every triplet is formed on the matrix cores
        "roofline": {"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                     "traffic": profiled_traffic(ISO_PROFILE, "k_triplets 3iso")},
    }


def batch_leg(e, steps, no_tune=False):
    """2ani over many randoms in one persistent launch: Neyman-Scott data
    (8192 parents x 32, sigma 8, N = 262,144, box 1000) against 64 uniform
    random catalogs with N_i = round(exp(U[ln 2e4, ln 4e5])), rmax 120,
    30 x 30 bins, no BPC: DD, 64 RR and 64 DR = 129 histograms per step.
    Reported: the batched step (plan reused, one graph launch) against the
    same 129 histograms as separate pcf_histogram calls."""
    d = synth.neyman_scott(8192, 32, 8.0, BOX, seed=21)
    rng = np.random.default_rng(22)
    sizes = [int(round(np.exp(rng.uniform(np.log(2e4), np.log(4e5))))) for _ in range(64)]
    cd = e.upload(d)
    crs = [e.upload(synth.uniform(n, BOX, seed=100 + i)) for i, n in enumerate(sizes)]
    jobs = [da.make_job(da.PCF_2ANI, da.PCF_XX, bins=30, dmax=120.0, cat_x=cd)]
    for cr in crs:
        jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XX, bins=30, dmax=120.0, cat_x=cr))
        jobs.append(da.make_job(da.PCF_2ANI, da.PCF_XY, bins=30, dmax=120.0, cat_x=cd, cat_y=cr))
    tune = None if no_tune else e.autotune(jobs)
    t0 = time.perf_counter()
    res = e.histogram_batch(jobs)  # plan: grids, work lists, graph capture
    t1 = time.perf_counter()
    first = e.stats()
    kms = 0.0
    evals = 0
    for _ in range(steps):
        res = e.histogram_batch(jobs)
        st = e.stats()
        assert st["graph_replay"] == 1
        kms += st["kernel_ms"]
        evals += st["evaluations"]
    t2 = time.perf_counter()
    useful = sum(int(c.sum()) for c, _ in res)
    single_ms = 0.0
    t3 = time.perf_counter()
    for k, j in enumerate(jobs):
        c, _ = e.histogram(j)
        single_ms += e.stats()["kernel_ms"]
        if not np.array_equal(c, res[k][0]):
            raise SystemExit(f"batched and single histograms differ (job {k})")
    t4 = time.perf_counter()
    # Every histogram of the batch against the plain FP64 path, bin by bin.
    worst = 0.0
    for k, j in enumerate(jobs):
        j64 = da.make_job(j.type, j.hist, bins=j.bins, dmax=j.dmax, cat_x=j.cat_x,
                          cat_y=j.cat_y, precision=da.PCF_PREC_FP64)
        name = "DD" if k == 0 else ("RR%d" if j.hist == da.PCF_XX else "DR%d") % ((k - 1) // 2)
        worst = max(worst, compare(f"2ani batch {name}", res[k], e.histogram(j64))["max_rel_err_w"])
    parity = {"histograms": len(jobs), "counts_equal": True, "max_rel_err_w": worst}
    e.release(cd)
    for cr in crs:
        e.release(cr)
    ms = (t2 - t1) * 1e3 / steps
    ach = evals * FLOPS_PER_PAIR / (kms * 1e-3)
    return {"workload": "2ani N=262144 Neyman-Scott (8192x32, sigma 8) vs 64 uniform randoms "
                        f"(N {min(sizes)}..{max(sizes)}, total {sum(sizes)}), box 1000, rmax 120, "
                        "30x30 bins, DD + 64 RR + 64 DR in one persistent launch",
            "value": useful / (ms * 1e-3), "unit": "pairs/s", "ms_per_step": ms, "steps": steps,
            "kernel_ms_per_step": kms / steps, "evaluations_per_step": evals // steps,
            "first_call_ms": (t1 - t0) * 1e3, "chunks": first["chunks"], "autotune": tune,
            "work_items": first["work_items"],
            "unbatched_ms": (t4 - t3) * 1e3, "unbatched_kernel_ms": single_ms,
            "speedup_vs_unbatched": (t4 - t3) * 1e3 / ms,
            "parity_vs_fp64": parity,
            "roofline": {"bound": "valu", "achieved": ach / 1e12, "peak": PEAK_FP32 / 1e12,
                         "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                         "traffic": profiled_traffic(BATCH_PROFILE, "k_pairs_mixed batch")}}


# Expected ordered triplets above which a 3iso request of the serving mix is
# capped: the mix's N (log-uniform up to 1,048,576) with rmax up to 120 in a box
# of 250 would ask for ~1e17 triplets (about a day of GPU time) in one
# request.  The cap lowers N until N (N V / L^3)^2 <= TRIPLET_CAP, V the
# sphere of radius rmax: at most ~1e10 triplet evaluations (~10 ms).
TRIPLET_CAP = 1e10


def serving_leg(requests=2000, rate=100.0, lanes=4, max_batch=8, pool_size=48, check=40):
    """Open-loop serving of the serving configuration: a Poisson stream of
    `requests` requests; each picks its type (2iso 50%, 2ani 30%, 3iso 20%),
    N log-uniform in [1024, 1048576], bins in {20, 30, 40}, rmax U[40, 120],
    box 250 or 1000, periodic 50%, uniform or Neyman-Scott data, all from its
    own seed.  The catalogs are drawn from a pool of `pool_size` resident
    catalogs (sizes log-uniform over the same range, uniform or Neyman-Scott,
    both boxes): a request's seed picks the pool catalog of its box nearest
    its N, so grids stay resident and are reused across requests (DR and RR
    of the same catalogs).  3iso requests are capped by TRIPLET_CAP.  The
    engine runs `lanes` concurrent lanes with HBM-budget admission and
    continuous batching of up to `max_batch` compatible pair requests
    (different bins and rmax allowed).  Latency = submit -> result ready.  A
    sample of `check` requests is recomputed alone and must match."""
    rng = np.random.default_rng(31)
    lo, hi = np.log(1024), np.log(1048576)
    pool = []  # (box, n, xyzw)
    for k in range(pool_size):
        box = 250.0 if k % 2 == 0 else 1000.0
        n = int(round(np.exp(lo + (hi - lo) * (k // 2 + 0.5) / (pool_size // 2))))
        if k % 4 < 2:
            xyzw = synth.uniform(n, box, seed=300 + k)
        else:
            xyzw = synth.neyman_scott(max(1, n // 32), 32, 5.0, box, seed=300 + k)
        pool.append((box, len(xyzw), xyzw))
    by_box = {b: sorted((i for i, (bb, _, _) in enumerate(pool) if bb == b),
                        key=lambda i: pool[i][1]) for b in (250.0, 1000.0)}

    def pick(box, n):
        cand = by_box[box]
        sizes = np.array([pool[i][1] for i in cand])
        return cand[int(np.argmin(np.abs(np.log(sizes) - np.log(n))))]

    jobs, capped = [], 0
    for q in range(requests):
        rq = np.random.default_rng(1000 + q)
        u = rq.random()
        type_ = da.PCF_2ISO if u < 0.5 else (da.PCF_2ANI if u < 0.8 else da.PCF_3ISO)
        n = float(np.exp(rq.uniform(lo, hi)))
        bins = int(rq.choice([20, 30, 40]))
        dmax = float(rq.uniform(40.0, 120.0))
        box = 250.0 if rq.random() < 0.5 else 1000.0
        bpc = bool(rq.random() < 0.5)
        cross = rq.random() < 0.5
        if type_ == da.PCF_3ISO:
            vol = 4.0 / 3.0 * np.pi * dmax ** 3
            nmax = (TRIPLET_CAP * box ** 6 / vol ** 2) ** (1.0 / 3.0)
            if n > nmax:
                n, capped = nmax, capped + 1
        x = pick(box, n)
        y = pick(box, n * float(np.exp(rq.uniform(-1.0, 1.0))))
        if type_ == da.PCF_3ISO:
            hist = da.PCF_XXY if cross else da.PCF_XXX
        else:
            hist = da.PCF_XY if cross else da.PCF_XX
        jobs.append((type_, hist, bins, dmax, box, bpc, x, y if cross else -1))
    with da.Engine(0) as e:
        ids = [e.upload(xyzw) for _, _, xyzw in pool]
        made = [da.make_job(t, h, bins=b, dmax=dm, cat_x=ids[x], cat_y=ids[y] if y >= 0 else -1,
                            bpc=bp, box=bx) for t, h, b, dm, bx, bp, x, y in jobs]
        e.set_hbm_budget(0, lanes)
        e.set_max_batch(max_batch)
        gaps = np.random.default_rng(32).exponential(1.0 / rate, requests)
        tickets = []
        t0 = time.perf_counter()
        t_next = t0
        for j, g in zip(made, gaps):
            t_next += g
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(e.submit(j))
        lat, queue, useful_pairs, useful_trip, batched = [], [], 0, 0, 0
        results = []
        for t, j in zip(tickets, made):
            c, w, st = e.wait(t, j.type, j.bins)
            lat.append(st["latency_ms"])
            queue.append(st["queue_ms"])
            batched += st["batched_jobs"] > 1
            if j.type in (da.PCF_2ISO, da.PCF_2ANI):
                useful_pairs += int(c.sum())
            else:
                useful_trip += int(c.sum())
            results.append((c, w))
        t1 = time.perf_counter()
        # Spot check: a sample of requests recomputed alone.
        srng = np.random.default_rng(33)
        for k in srng.choice(requests, size=min(check, requests), replace=False):
            c1, w1 = e.histogram(made[k])
            c, w = results[k]
            if not np.array_equal(c, c1) or not np.all(np.abs(w - w1) <= 1e-12 * np.abs(w1)):
                raise SystemExit(f"serving: request {k} differs from the same request alone")
    lat = np.array(lat)
    types = [t for t, *_ in jobs]
    return {"workload": f"{requests} Poisson requests at {rate:g}/s: 2iso/2ani/3iso "
                        f"{types.count(da.PCF_2ISO)}/{types.count(da.PCF_2ANI)}/"
                        f"{types.count(da.PCF_3ISO)}, N log-uniform [1024, 1048576] over a pool "
                        f"of {pool_size} resident catalogs (uniform or Neyman-Scott, box 250 or "
                        f"1000), bins 20/30/40, rmax U[40,120], 50% periodic, 50% cross; "
                        f"{capped} 3iso requests capped at {TRIPLET_CAP:g} expected triplets; "
                        f"{lanes} lanes, batching up to {max_batch}",
            "requests": requests, "wall_s": t1 - t0, "requests_per_s": requests / (t1 - t0),
            "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "mean_queue_ms": float(np.mean(queue)), "batched_requests": batched,
            "capped_3iso_requests": capped,
            "useful_pairs_per_s": useful_pairs / (t1 - t0),
            "useful_triplets_per_s": useful_trip / (t1 - t0),
            "checked_alone": min(check, requests)}


def analytic_leg(e):
    """3iso -bpc -a on the config-2 data catalog: fine periodic pair histograms
    (3,000 and 600,000 bins), ff_av / ff_av_ref and the analytic RRR / DDR bin
    cubes with the 200^3 sub-cube refinement (pcf_analytic_3iso)."""
    d = synth.neyman_scott(1024, 32, 5.0, 250.0, seed=11, weights="uniform")
    cd = e.upload(d)
    e.analytic_3iso(cd, bins=30, dmax=80.0, box=250.0)
    t0 = time.perf_counter()
    e.analytic_3iso(cd, bins=30, dmax=80.0, box=250.0)
    t1 = time.perf_counter()
    st = e.stats()
    e.release(cd)
    return {"workload": "3iso -bpc -a, N=32768 Neyman-Scott, box 250, rmax 80, 30 bins, "
                        "ptt 100, bn_ref 200 (DDR = DRR and RRR; DDD is the triplet leg)",
            "ms": (t1 - t0) * 1e3, "kernel_ms": st["kernel_ms"]}


def copy_probe():
    """A plain device copy of known size, the calibration run for the HBM byte
    counters (FETCH_SIZE / WRITE_SIZE per dispatch)."""
    import torch
    n = 1 << 28  # 256 Mi float32 = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}),
          flush=True)


def cpu_baseline(d, r):
    """The oracle (a C++/OpenMP restatement of the reference's algorithm) timed
    on the host cores, on the same NS step the GPU runs (DD + RR + DR of the
    two 1,048,576-point catalogs): in its FP64 canonical mode (the GPU path's
    arithmetic) and in its cpp32 mode (the reference CPU programs' FP32
    arithmetic and unordered-pair loop, cpp/2PCFisotropic/normal/2PCF.h:151-305).
    Threads: the box's CPU share (16 for one GPU) or fewer; the reference CPU
    programs' hard-wired 2 threads are timed on the DD of the data points in
    [0,500)^3."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    nproc = os.cpu_count() or 1
    cores = min(nproc, 16)
    t = time.perf_counter()
    tot = 0
    for a, b in ((d, None), (r, None), (d, r)):
        c, _ = O.pairs(a, b, bins=BINS, dmax=RMAX, threads=cores)
        tot += int(c.sum())
    dt = time.perf_counter() - t
    sel = np.all(d[:, :3] < 500.0, axis=1)
    sample = np.ascontiguousarray(d[sel])
    t = time.perf_counter()
    tot32 = float(O.cpp32_2iso(sample, None, bins=BINS, dmax=RMAX, threads=cores).sum())
    dt32 = time.perf_counter() - t
    t = time.perf_counter()
    c2, _ = O.pairs(sample, None, bins=BINS, dmax=RMAX, threads=2)
    dt2 = time.perf_counter() - t
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": tot / dt, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"the full NS step (DD + RR + DR, 2 x 1048576 points, rmax {RMAX}, "
                      f"{BINS} bins) in the oracle's FP64 grid mode: {dt:.2f} s",
            "algorithm": "oracle/pcf_oracle.cpp cell grid, ordered pairs, FP64 distances",
            "cpp32_value": tot32 / dt32,
            "cpp32_algorithm": "oracle cpp32 mode: the reference CPU program's FP32 "
                               "arithmetic and unordered-pair loop, same cell grid",
            "cpp32_sample": f"DD of the {len(sample)} data points in [0,500)^3, "
                            f"{cores} threads, {dt32:.2f} s",
            "nproc": nproc,
            "threads_note": "threads capped at the GPU box's CPU share of 16",
            "value_2_threads": float(c2.sum()) / dt2,
            "sample_2_threads": f"oracle FP64 DD of the {len(sample)} data points in "
                                f"[0,500)^3, {dt2:.2f} s",
            "cpu_model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--triplet-steps", type=int, default=1)
    ap.add_argument("--no-triplets", action="store_true")
    ap.add_argument("--no-3iso", action="store_true", help="skip the 3iso triplet leg only")
    ap.add_argument("--no-3ani", action="store_true", help="skip the 3ani triplet leg only")
    ap.add_argument("--no-analytic", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-rate", type=float, default=100.0,
                    help="Poisson arrival rate of the serving leg, requests/s")
    ap.add_argument("--serving-max-batch", type=int, default=8,
                    help="most queued pair requests one serving launch may take")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default launch configuration)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the 1 GiB device copy used to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs the pair counting on a single "
              "device (multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)

    d = synth.uniform(N, BOX, seed=1)
    r = synth.uniform(N, BOX, seed=2)
    with da.Engine(0) as e:
        cd, cr = e.upload(d), e.upload(r)
        jobs = ns_jobs(cd, cr)
        # One-time per workload (cached on the context, like the resident
        # grids): pair kernel form, cell size, LDS layout and occupancy.
        tuning = None if args.no_tune else e.autotune(jobs)
        for _ in range(args.warmup):
            run_batched_step(e, jobs)
        t0 = time.perf_counter()
        pairs = evals = rech = replays = 0
        kms = 0.0
        for _ in range(args.steps):
            res, st = run_batched_step(e, jobs)
            pairs += sum(int(c.sum()) for c, _ in res)
            evals += st["evaluations"]
            kms += st["kernel_ms"]
            rech += st["rechecks"]
            replays += st["graph_replay"]
        t1 = time.perf_counter()
        ms_step = (t1 - t0) * 1e3 / args.steps
        value = pairs / (t1 - t0)

        # Plain FP64 path, same step (grid build + DD + RR + DR).
        run_step(e, cd, cr, da.PCF_PREC_FP64)
        t2 = time.perf_counter()
        p64 = 0
        k64 = 0.0
        for _ in range(args.fp64_steps):
            p, _, k, _ = run_step(e, cd, cr, da.PCF_PREC_FP64)
            p64 += p
            k64 += k
        t3 = time.perf_counter()
        ms_step64 = (t3 - t2) * 1e3 / args.fp64_steps
        ns_parity = {n: compare(f"NS {n}", a, b)
                     for n, a, b in zip(("DD", "RR", "DR"), res,
                                        run_step_hists(e, cd, cr, da.PCF_PREC_FP64))}
        trip = None if args.no_triplets or args.no_3iso else triplet_leg(e, args.triplet_steps)
        trip_ani = (None if args.no_triplets or args.no_3ani
                    else triplet_ani_leg(e, max(1, args.triplet_steps)))
        ana = None if args.no_analytic else analytic_leg(e)
        bat = None if args.no_batch else batch_leg(e, args.batch_steps, args.no_tune)

    serving = None if args.no_serving else serving_leg(args.serving_requests,
                                                       rate=args.serving_rate,
                                                       max_batch=args.serving_max_batch)
    achieved = evals * FLOPS_PER_PAIR / (kms * 1e-3)
    # The dominant kernel: the Gram form forms d2 with four chained FMAs on the
    # matrix cores (FP32 MFMA peak 157.3 TFLOP/s, the same figure as the packed
    # FP32 vector peak); the direct form runs on the vector units.
    # Untuned, the library's automatic choice is the Gram form for this unweighted 2iso step.
    kern = da.PCF_PAIR_GRAM if tuning is None else tuning.get("pair_kernel", da.PCF_PAIR_GRAM)
    if kern == da.PCF_PAIR_AUTO:
        kern = da.PCF_PAIR_GRAM
    bound = "mfma" if kern == da.PCF_PAIR_GRAM else "valu"
    prof = PAIR_PROFILE.get(kern)
    out = {
        "metric": "pairs/s (2iso DD+RR+DR, useful pairs)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-recheck",
        "data": "synthetic (uniform, seeded)",
        "config": {"workload": "2iso N=1048576 data + 1048576 random, box 1000, rmax 80, "
                               "30 bins, DD+RR+DR in one batched launch (grid build "
                               "included), no BPC",
                   "n_data": N, "n_random": N, "box": BOX, "rmax": RMAX, "bins": BINS},
        "roofline": {"bound": bound, "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": profiled_traffic(prof, "k_pairs_mixed NS")},
        "pair_kernel_ms_per_step": kms / args.steps,
        "evaluations_per_step": evals // args.steps,
        "candidates_per_useful_pair": evals / max(pairs, 1),
        "rechecks_per_step": rech // args.steps,
        "graph_replays": replays,
        "parity_vs_fp64": ns_parity,
        "fp64_plain_ms_per_step": ms_step64,
        "fp64_plain_kernel_ms_per_step": k64 / args.fp64_steps,
        "speedup_vs_fp64_plain": ms_step64 / ms_step,
    }
    if tuning is not None:
        out["autotune"] = tuning
    if trip is not None:
        out["triplets"] = trip
    if trip_ani is not None:
        out["triplets_3ani"] = trip_ani
    if ana is not None:
        out["analytic_3iso"] = ana
    if bat is not None:
        out["batched_2ani_64_randoms"] = bat
    if serving is not None:
        out["serving"] = serving
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(d, r)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
