"""A conversion service on one GPU: jobs arrive, are admitted by HBM budget,
and are converted device-resident.

Each job is a synthetic LAS tile of the 500 m generator (LasGen preset 0)
with its own seed and a point count drawn log-uniformly from
[min_points, max_points]; jobs arrive by a Poisson process of `rate` jobs
per second (exponential gaps, seeded).  A job is admitted when its records
fit the HBM budget beside the jobs already admitted and the library's
working memory for the largest admitted job; admission places its records
in HBM (the generator writes them there, standing in for the upload).
Admitted jobs are converted one at a time in arrival order on the
library's stream (doChunking + doIndexing to the octree image in HBM,
pcg_convert_device); while one converts, the next jobs keep arriving on
the wall clock.  Latency is completion minus arrival.

No reference counterpart (the reference converter is a one-shot process).
"""
import ctypes
import math
import random
import time

from . import LasGen, _check, lib

# device memory of a conversion beyond its input records, per point (the
# library's chunk records, octree-build buffers, sampling arena and scratch
# and octree image at 200M points: about 485 bytes per point)
WORK_BYTES_PER_POINT = 520
RECORD_BYTES = 26  # LAS point format 2, the generator's records


class Job:
    __slots__ = ("id", "seed", "points", "arrival", "admitted", "started", "done", "d_las", "header", "stats",
                 "bytes")

    def __init__(self, jid, seed, points, arrival):
        self.id, self.seed, self.points, self.arrival = jid, seed, points, arrival
        self.admitted = self.started = self.done = None
        self.d_las = None
        self.header = None
        self.stats = None
        self.bytes = 0


def make_jobs(num_jobs, rate, min_points, max_points, seed):
    """Arrival times (s, from 0) and sizes of the job stream."""
    rng = random.Random(seed)
    t, jobs = 0.0, []
    lo, hi = math.log(min_points), math.log(max_points)
    for i in range(num_jobs):
        t += rng.expovariate(rate)
        n = int(round(math.exp(rng.uniform(lo, hi))))
        jobs.append(Job(i, seed * 1_000_003 + i, n, t))
    return jobs


def percentile(values, q):
    v = sorted(values)
    if not v:
        return None
    k = (len(v) - 1) * q / 100.0
    f = math.floor(k)
    c = min(f + 1, len(v) - 1)
    return v[f] + (v[c] - v[f]) * (k - f)


def serve(num_jobs=200, rate=20.0, min_points=100_000, max_points=50_000_000, seed=0, method="poisson",
          hbm_budget=None, log=None):
    """Runs the job stream in real time; returns a summary dict."""
    L = lib()
    free, total = ctypes.c_uint64(), ctypes.c_uint64()
    _check(L.pcg_device_memory(ctypes.byref(free), ctypes.byref(total)))
    if hbm_budget is None:
        hbm_budget = int(0.85 * total.value)
    jobs = make_jobs(num_jobs, rate, min_points, max_points, seed)
    pending = list(jobs)  # not yet arrived, by arrival time
    queue = []            # arrived, waiting for admission (FIFO)
    admitted = []         # records in HBM, waiting for conversion (FIFO)
    resident = 0          # input bytes of admitted jobs
    waits = 0
    stats = (ctypes.c_double * 24)()
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    def work_bytes(n):
        return n * WORK_BYTES_PER_POINT

    def admit(job):
        nonlocal resident
        g = LasGen(0, job.points, seed=job.seed)
        rl = g.record_length
        d = ctypes.c_void_p()
        _check(L.pcg_device_alloc(ctypes.byref(d), job.points * rl + 64))
        q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
        dq = ctypes.c_void_p()
        _check(L.pcg_device_alloc(ctypes.byref(dq), 64))
        _check(L.pcg_copy(dq, q, 24, 0))
        _check(L.pcg_lasgen_device(g.buf, 0, job.points, d, dq))
        _check(L.pcg_copy(q, dq, 24, 1))
        L.pcg_device_free(dq)
        header = ctypes.create_string_buffer(375)
        _check(L.pcg_lasgen_header(g.buf, q, header))
        job.d_las, job.header = d, header
        job.admitted = now()
        job.bytes = job.points * rl
        resident += job.bytes

    while pending or queue or admitted:
        t = now()
        while pending and pending[0].arrival <= t:
            queue.append(pending.pop(0))
        # admission: FIFO while the budget holds the inputs plus the working
        # memory of the largest admitted job
        while queue:
            j = queue[0]
            biggest = max([a.points for a in admitted] + [j.points])
            if resident + j.points * RECORD_BYTES + work_bytes(biggest) > hbm_budget and admitted:
                waits += 1
                break
            admit(queue.pop(0))
            admitted.append(j)
        if not admitted:
            if pending:
                time.sleep(max(0.0, pending[0].arrival - now()))
            continue
        j = admitted.pop(0)
        j.started = now()
        _check(L.pcg_convert_device(j.d_las, j.points, j.header, method.encode(), None, stats))
        j.done = now()
        j.stats = (stats[2], stats[3], stats[7])
        L.pcg_device_free(j.d_las)
        resident -= j.bytes
        j.d_las = None
        if log:
            log(f"job {j.id}: {j.points} pts, wait {j.started - j.arrival:.3f} s, "
                f"convert {j.done - j.started:.3f} s")
    wall = now()
    lat = [j.done - j.arrival for j in jobs]
    conv = [j.done - j.started for j in jobs]
    pts = sum(j.points for j in jobs)
    busy = sum(conv)
    return {
        "jobs": num_jobs,
        "rate_jobs_per_s": rate,
        "points": pts,
        "wall_s": round(wall, 3),
        "aggregate_Mpts_per_s": round(pts / wall / 1e6, 2),
        "busy_Mpts_per_s": round(pts / busy / 1e6, 2),
        "gpu_busy_fraction": round(busy / wall, 3),
        "latency_p50_s": round(percentile(lat, 50), 4),
        "latency_p99_s": round(percentile(lat, 99), 4),
        "latency_max_s": round(max(lat), 4),
        "convert_p50_s": round(percentile(conv, 50), 4),
        "convert_p99_s": round(percentile(conv, 99), 4),
        "admission_waits": waits,
        "octree_bytes_conserved": all(j.stats[2] == j.points * (RECORD_BYTES + 1) for j in jobs)
        if method == "poisson" else None,
        "hbm_budget_bytes": hbm_budget,
        "sizes": [min_points, max_points],
    }
