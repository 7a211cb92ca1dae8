#!/usr/bin/env python3
"""Throughput of the gfx950 counting-sort chunker + octree indexer + Poisson sampler.

One step = doChunking + doIndexing (Poisson) of a seeded synthetic LAS point
set that is already resident in HBM (5 km x 5 km tile in UTM-like
coordinates, LAS 1.4 point format 2, scale 0.001) up to the octree.bin image
in HBM plus hierarchy.bin / metadata.json on the host.  No files are written
inside the timed region.

--config c2: one step = doIndexing (default -m poisson_average) of 4096
pre-chunked level-4 chunks resident in HBM, log-normal point counts (median
50 000, clipped to 1 000 .. 5 000 000; about 400M points), 34-byte records
with three extra-bytes attributes (potreeconverter_amd.ChunkGen).

Prints one JSON line (see README for the fields).  `cpu_baseline` times the
single-threaded CPU oracle (a restatement of the reference algorithm) on a
bounded prefix of the same point stream, on this machine's host.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
STATS = 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", default="c1", choices=["c1", "c2", "c3", "c4"],
                    help="c1: 200M-point LAS tile, chunking + indexing; c2: indexing of 4096 pre-chunked chunks; "
                         "c3: the precision scene (offset 1e6 m, scale 0.0001, near-duplicate clusters, lattices "
                         "at the node spacing) at 1B points, chunking + indexing; c4: a serving loop of "
                         "Poisson-arriving conversion jobs (potreeconverter_amd.serve)")
    ap.add_argument("--jobs", type=int, default=200, help="c4: number of jobs")
    ap.add_argument("--rate", type=float, default=20.0, help="c4: mean arrivals per second")
    ap.add_argument("--c2-plain-median", type=int, default=2_500,
                    help="c2: median chunk size of the batched-vs-plain comparison (0 skips it)")
    ap.add_argument("--sample-groups", type=int, default=None,
                    help="library option sample_groups (streams of the chunk-subtree sampling; default: the library's)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="library option (pcg_set_option) set before the run; repeatable")
    ap.add_argument("--autotune", action="store_true",
                    help="c1/c3: search the sampling launch options on the workload before the timed steps")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=None, help="default: 200M (c1), 1B (c3)")
    ap.add_argument("--method", default=None, choices=["poisson", "poisson_average", "random"],
                    help="default: poisson (c1), poisson_average (c2)")
    ap.add_argument("--encoding", default="DEFAULT",
                    help="--encoding of the converter: DEFAULT (octree image in HBM) or BROTLI (compressed nodes "
                         "assembled in host memory)")
    ap.add_argument("--random-seed", type=int, default=0, help="seed of -m random's leaf shuffles")
    ap.add_argument("--cpu-points", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-io", action="store_true", help="skip the upload / octree write timings")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-points", type=int, default=20_000_000,
                    help="size of the batched-vs-plain sampling comparison (0 skips it)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="one device-to-device copy of this many MiB before the steps (HBM counter calibration)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(points, seed, method, preset=1, total=200_000_000):
    """Times the CPU side on the first `points` records of the same stream:
    doChunking by the oracle (one thread; the reference chunker cannot be
    built here: it needs LASzip's CMake-generated header) and doIndexing by
    the reference indexer itself (oracle/_ref/ref_indexer, built from the
    reference sources), with its own task pool of hardware_concurrency + 4
    threads over the chunks (indexer.cpp:1638-1707).  Without the reference
    build the oracle's single-threaded indexer is timed instead."""
    import potreeconverter_amd as pc

    ocl = os.path.join(HERE, "oracle", "_build", "liboracle.so")
    if not os.path.exists(ocl):
        return None
    O = ctypes.CDLL(ocl)
    O.orc_convert.argtypes = [ctypes.c_char_p] * 3 + [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    ref = os.path.join(HERE, "oracle", "_ref", "ref_indexer")
    g = pc.LasGen(preset, total, seed=seed)
    g.n = points
    cpus = len(os.sched_getaffinity(0))
    hw = os.cpu_count() or cpus
    # generate the prefix of the full stream and write it as its own LAS file
    with tempfile.TemporaryDirectory() as td:
        las = os.path.join(td, "sample.las")
        ctypes.c_int64.from_buffer(g.buf, 8).value = points  # LasGenParams.n (tiling stays that of the full set)
        # keep the full set's tiling: tiles_per_side / num_tiles were fixed by the preset call
        pc.lib().pcg_lasgen_write(g.buf, las.encode())
        t = (ctypes.c_double * 8)()
        out = os.path.join(td, "out")
        if os.path.exists(ref):
            rc = O.orc_convert(las.encode(), out.encode(), method.encode(), 1, 1, t)
            if rc != 0:
                return None
            chunk_s = t[2]
            t0 = time.time()
            r = subprocess.run([ref, out, method, "sample", str(points), "DEFAULT"], stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL)
            index_s = time.time() - t0
            if r.returncode != 0:
                return None
            kind, cores = "reference", cpus
            how = (f"doChunking by the oracle on 1 thread {chunk_s:.2f}s + doIndexing by the reference indexer "
                   f"{index_s:.2f}s ({hw + 4} threads on {cpus} available CPUs, incl. its chunk file reads and "
                   f"octree.bin / hierarchy writes)")
        else:
            t0 = time.time()
            rc = O.orc_convert(las.encode(), out.encode(), method.encode(), 0, 0, t)
            if rc != 0:
                return None
            chunk_s, index_s = t[2], t[3]
            kind, cores = "port", 1
            how = f"oracle on 1 thread: chunking {chunk_s:.2f}s + indexing {index_s:.2f}s"
    return {
        "value": points / (chunk_s + index_s) / 1e6,
        "unit": "Mpts/s",
        "cores": cores,
        "kind": kind,
        "sample": f"first {points} points of the same generated stream; " + how,
        "cpu_model": cpu_model(),
        "parts_s": {"chunking": round(chunk_s, 3), "indexing": round(index_s, 3)},
        "chunking_Mpts_per_s_1_thread": round(points / chunk_s / 1e6, 2),
        "indexing_Mpts_per_s": round(points / index_s / 1e6, 2),
    }


def io_times(L, d_las, n, rl, header, method):
    """PCIe and file I/O around one conversion, outside the timed steps:
    the LAS records host -> device (pageable host memory), and the octree
    written to disk (octree.bin downloaded from HBM + hierarchy.bin +
    metadata.json) as the extra wall time of a conversion with an output
    directory over one without."""
    import shutil

    import potreeconverter_amd as pc

    out = {}
    host = ctypes.create_string_buffer(n * rl)
    pc._check(L.pcg_copy(host, d_las, n * rl, 1))
    t0 = time.perf_counter()
    pc._check(L.pcg_copy(d_las, host, n * rl, 0))
    out["upload_s"] = round(time.perf_counter() - t0, 4)
    out["upload_GBps"] = round(n * rl / out["upload_s"] / 1e9, 2)
    del host
    stats = (ctypes.c_double * STATS)()
    with tempfile.TemporaryDirectory() as td:
        if shutil.disk_usage(td).free < 3 * n * (rl + 1):
            return out
        t0 = time.perf_counter()
        pc._check(L.pcg_convert_device(d_las, n, header, method.encode(), None, stats))
        plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        pc._check(L.pcg_convert_device(d_las, n, header, method.encode(), os.path.join(td, "o").encode(), stats))
        written = time.perf_counter() - t0
        out["download_and_write_s"] = round(written - plain, 4)
        out["octree_bytes"] = int(stats[7])
    return out


def plain_vs_batched(points, seed, method):
    """Node sampling of the same input with every node of a level in one
    launch per kernel (batched) and with one launch per node (plain)."""
    import potreeconverter_amd as pc

    L = pc.lib()
    g = pc.LasGen(1, 200_000_000, seed=seed)  # the prefix of the benchmark's point stream
    ctypes.c_int64.from_buffer(g.buf, 8).value = points
    rl = g.record_length
    d_las, d_q = ctypes.c_void_p(), ctypes.c_void_p()
    pc._check(L.pcg_device_alloc(ctypes.byref(d_las), points * rl + 64))
    pc._check(L.pcg_device_alloc(ctypes.byref(d_q), 64))
    qmm = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    pc._check(L.pcg_copy(d_q, qmm, 24, 0))
    pc._check(L.pcg_lasgen_device(g.buf, 0, points, d_las, d_q))
    pc._check(L.pcg_copy(qmm, d_q, 24, 1))
    header = ctypes.create_string_buffer(375)
    pc._check(L.pcg_lasgen_header(g.buf, qmm, header))
    st = {}
    try:
        for mode in ("batched", "batched", "plain"):
            pc._check(L.pcg_set_option(b"plain_sampling", 1 if mode == "plain" else 0))
            stats = (ctypes.c_double * STATS)()
            pc._check(L.pcg_convert_device(d_las, points, header, method.encode(), None, stats))
            st[mode] = list(stats)
    finally:
        L.pcg_set_option(b"plain_sampling", 0)
        L.pcg_device_free(d_las)
        L.pcg_device_free(d_q)
    b, p = st["batched"], st["plain"]
    cand = b[10]
    return {
        "points": points,
        "candidates": int(cand),
        "sampled_nodes": int(b[9]),
        "batched_sampling_s": round(b[5], 4),
        "plain_sampling_s": round(p[5], 4),
        "batched_Mcand_per_s": round(cand / b[5] / 1e6, 1),
        "plain_Mcand_per_s": round(cand / p[5] / 1e6, 1),
        "speedup": round(p[5] / b[5], 2),
        "same_output_bytes": b[7] == p[7],
    }


C2 = dict(num_chunks=4096, median=50_000, sigma=1.16, min_count=1_000, max_count=5_000_000)


def c2_indexing(args):
    """Generates the pre-chunked set in HBM, then times doIndexing on it."""
    import potreeconverter_amd as pc

    L = pc.lib()

    def setup(median):
        g = pc.ChunkGen(**dict(C2, median=median), seed=args.seed)
        d = ctypes.c_void_p()
        pc._check(L.pcg_device_alloc(ctypes.byref(d), g.total * g.RECORD_BYTES + 64))
        meta = g.device(d.value)
        n = len(g.names)
        names = (ctypes.c_char_p * n)(*[x.encode() for x in g.names])
        counts = (ctypes.c_int64 * n)(*g.counts)
        return g, d, meta.encode(), names, counts

    stats = (ctypes.c_double * STATS)()

    def index(job):
        g, d, meta, names, counts = job
        pc._check(L.pcg_do_indexing_device(d, len(g.names), names, counts, meta, args.method.encode(),
                                           args.encoding.encode(), b"c2", None, stats))
        return list(stats)

    job = setup(C2["median"])
    g = job[0]
    for _ in range(args.warmup):
        index(job)
    pc._check(L.pcg_synchronize())
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per.append(index(job))
    pc._check(L.pcg_synchronize())
    elapsed = time.perf_counter() - t0
    L.pcg_device_free(job[1])
    last = per[-1]
    samp_s = sum(p[5] for p in per) / len(per)
    bpp = g.RECORD_BYTES
    cand = last[10]
    # node sampling: 12 + 2 * bpp bytes per candidate (distance key + index,
    # the record read once to gather and once to route)
    samp_bytes = cand * (12 + 2 * bpp)
    achieved = samp_bytes / samp_s / 1e9 if samp_s > 0 else 0.0
    plain = None
    if args.c2_plain_median > 0:
        small = setup(args.c2_plain_median)
        st = {}
        try:
            for mode in ("batched", "batched", "plain"):
                pc._check(L.pcg_set_option(b"plain_sampling", 1 if mode == "plain" else 0))
                st[mode] = index(small)
        finally:
            L.pcg_set_option(b"plain_sampling", 0)
            L.pcg_device_free(small[1])
        b, p = st["batched"], st["plain"]
        plain = {
            "points": small[0].total, "median": args.c2_plain_median, "candidates": int(b[10]),
            "sampled_nodes": int(b[9]), "batched_sampling_s": round(b[5], 4), "plain_sampling_s": round(p[5], 4),
            "batched_Mcand_per_s": round(b[10] / b[5] / 1e6, 1), "plain_Mcand_per_s": round(b[10] / p[5] / 1e6, 1),
            "speedup": round(p[5] / b[5], 2), "same_output_bytes": b[7] == p[7],
        }
    line = {
        "metric": "Mpts/s doIndexing of 4096 pre-chunked log-normal chunks (octree + sampling), batched vs plain",
        "value": round(g.total / (elapsed / args.steps) / 1e6, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"4096 level-4 chunks of a 2048 m cube, log-normal counts (median 50000, sigma 1.16, "
                        f"clip 1000..5000000), {g.total} pts, PDRF2 + 3 extra bytes (34 B), seed {args.seed}",
            "points": g.total, "chunks": len(g.names), "method": args.method, "encoding": args.encoding,
            "path": "batched (all nodes of a level per launch)",
        },
        "phases_s": {
            "indexing": round(last[3], 4), "octree_build": round(last[4], 4), "sampling": round(last[5], 4),
            "octree_assembly": round(last[6], 4), "host_tree_setup": round(last[21], 4),
            "host_completion_order": round(last[22], 4), "host_hierarchy_metadata": round(last[23], 4),
        },
        "counters": {
            "nodes": int(last[8]), "sampled_nodes": int(last[9]), "candidates": int(cand),
            "fp32_pair_tests": int(last[12]), "fp64_rechecks": int(last[11]), "max_fp32_rel_err": last[13],
            "libstdcxx_order_nodes": int(last[14]), "sequential_nodes": int(last[15]),
            "octree_rounds": int(last[17]), "depth": int(last[18]), "octree_bytes": int(last[7]),
        },
        "roofline": {
            "bound": "hbm", "kernel": "node sampling (all sampling kernels, sampling phase)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": samp_bytes,
            "kernel_ms": round(samp_s * 1e3, 3),
        },
        "plain_vs_batched": plain,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def c4_serving(args):
    """Poisson arrivals of conversion jobs (100k .. 50M points, log-uniform),
    HBM-budget admission, device-resident conversion; one JSON line."""
    from potreeconverter_amd import serve

    # a short stream first: code paths and buffers warm
    serve.serve(num_jobs=max(args.warmup, 0) * 4, rate=args.rate, min_points=100_000, max_points=2_000_000,
                seed=args.seed + 1, method=args.method)
    r = serve.serve(num_jobs=args.jobs, rate=args.rate, min_points=100_000, max_points=50_000_000, seed=args.seed,
                    method=args.method)
    line = {
        "metric": "serving loop: aggregate Mpts/s and p50/p99 job latency (Poisson arrivals, HBM admission)",
        "value": r["aggregate_Mpts_per_s"],
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.jobs,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] / args.jobs * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.jobs} jobs, Poisson arrivals at {args.rate}/s, 500 m LAS tiles (preset 0, "
                               f"per-job seeds) of 100k..50M points log-uniform, -m {args.method}",
                   "jobs": args.jobs, "rate": args.rate, "method": args.method},
        "serving": r,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md, "
                                   "--gpus above 1 is not supported"}))
        sys.exit(2)
    if args.method is None:
        args.method = "poisson_average" if args.config == "c2" else "poisson"
    if args.sample_groups is not None:
        import potreeconverter_amd as pc
        pc._check(pc.lib().pcg_set_option(b"sample_groups", args.sample_groups))
    for kv in args.option:
        import potreeconverter_amd as pc
        key, value = kv.split("=", 1)
        pc._check(pc.lib().pcg_set_option(key.encode(), int(value)))
    if args.config == "c2":
        c2_indexing(args)
        return
    if args.config == "c4":
        c4_serving(args)
        return
    import potreeconverter_amd as pc

    L = pc.lib()
    c3 = args.config == "c3"
    preset = 2 if c3 else 1
    n = args.points or (1_000_000_000 if c3 else 200_000_000)
    if c3 and args.plain_points == 20_000_000:
        args.plain_points = 0  # the per-node plain path is measured on the c1 workload
    g = pc.LasGen(preset, n, seed=args.seed)
    rl = g.record_length
    d_las = ctypes.c_void_p()
    pc._check(L.pcg_device_alloc(ctypes.byref(d_las), n * rl + 64))
    qmm = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    d_q = ctypes.c_void_p()
    pc._check(L.pcg_device_alloc(ctypes.byref(d_q), 64))
    pc._check(L.pcg_copy(d_q, qmm, 24, 0))
    pc._check(L.pcg_lasgen_device(g.buf, 0, n, d_las, d_q))
    pc._check(L.pcg_copy(qmm, d_q, 24, 1))
    header = ctypes.create_string_buffer(375)
    pc._check(L.pcg_lasgen_header(g.buf, qmm, header))
    stats = (ctypes.c_double * STATS)()

    if args.copy_calibration_mib > 0:
        cb = args.copy_calibration_mib << 20
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        pc._check(L.pcg_device_alloc(ctypes.byref(src), cb))
        pc._check(L.pcg_device_alloc(ctypes.byref(dst), cb))
        pc._check(L.pcg_copy(dst, src, cb, 2))
        pc._check(L.pcg_synchronize())
        L.pcg_device_free(src)
        L.pcg_device_free(dst)

    if args.method == "random":
        pc._check(L.pcg_set_option(b"random_seed", args.random_seed))

    def step():
        if args.encoding == "DEFAULT":
            pc._check(L.pcg_convert_device(d_las, n, header, args.method.encode(), None, stats))
        else:
            pc._check(L.pcg_convert_device_encoded(d_las, n, header, args.method.encode(), args.encoding.encode(),
                                                   None, stats))

    tuned = None
    if args.autotune:
        # the launch options searched on this workload before the timed steps
        chosen, best_s, default_s = pc.autotune_device(d_las.value, n, header, args.method, passes=2, repeats=2)
        tuned = {"options": chosen, "best_ms": round(best_s * 1e3, 3), "start_ms": round(default_s * 1e3, 3)}
    for _ in range(args.warmup):
        step()
    pc._check(L.pcg_synchronize())
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        per.append(list(stats))
    pc._check(L.pcg_synchronize())
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    value = n / (elapsed / args.steps) / 1e6

    # roofline of the chunking pass (count + distribute): 79 B/pt algorithmic
    # bytes for 26-byte input records and 27-byte output records
    bpp_in, bpp_out = rl, rl + 1
    chunk_bytes = n * (2 * bpp_in + bpp_out)
    chunk_s = sum(p[2] for p in per) / len(per)
    kern = {
        "count_ms": sum(p[19] for p in per) / len(per),
        "distribute_ms": sum(p[20] for p in per) / len(per),
    }
    kern_s = (kern["count_ms"] + kern["distribute_ms"]) / 1e3
    achieved = chunk_bytes / kern_s / 1e9 if kern_s > 0 else 0.0
    last = per[-1]
    # HBM bytes per launch of the chunking pass from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE runs of this workload (copy-calibrated)
    traffic = None
    tpath = os.path.join(HERE, "profiles", "chunking_200m_hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("points") == n and args.seed == 0:
            traffic = tr["traffic_bytes"]
    plain = None
    if args.plain_points > 0:
        plain = plain_vs_batched(min(args.plain_points, n), args.seed, args.method)
    cpu = None if args.no_cpu else cpu_baseline(min(args.cpu_points, n), args.seed, args.method, preset, n)
    io = None if (args.no_io or c3) else io_times(L, d_las, n, rl, header, args.method)
    workload = (f"{n} pts synthetic LAS PDRF2 5x5 km (offset 500000,5000000,0; scale 0.001), "
                f"terrain fBm + 20 buildings, seed {args.seed}")
    metric = "Mpts/s count-sort+octree+Poisson at 200M pts, batched vs plain; % HBM peak"
    if c3:
        workload = (f"{n} pts precision scene PDRF2 (offset 1e6 m, scale 0.0001; 10% near-duplicate clusters "
                    f"jittered +-1 unit, 5% on a lattice whose pitch is the level-4 node spacing), seed {args.seed}")
        metric = "Mpts/s count-sort+octree+Poisson at 1B pts (precision scene); FP64 re-check fraction"
    precision = None
    if c3 or args.method != "random":
        keys = (8, 9, 10, 11, 12, 14, 15, 7, 17, 18)
        precision = {
            "fp64_rechecks": int(last[11]), "fp32_pair_tests": int(last[12]),
            "recheck_fraction": last[11] / last[12] if last[12] else None,
            "max_fp32_rel_err": last[13],
            "points_conserved": (int(last[7]) == n * bpp_out) if args.method == "poisson" else None,
            "counters_equal_over_steps": all(all(p[k] == per[0][k] for k in keys) for p in per),
        }
    line = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "points": n,
            "method": args.method,
            "encoding": args.encoding,
            "path": "batched (all nodes of a level per launch)",
            "grid": 128 if n < 100_000_000 else (256 if n < 500_000_000 else 512),
        },
        "phases_s": {
            "chunking": round(last[2], 4), "count": round(last[0], 4), "lut+distribute": round(last[1], 4),
            "indexing": round(last[3], 4), "octree_build": round(last[4], 4), "sampling": round(last[5], 4),
            "octree_assembly": round(last[6], 4), "host_tree_setup": round(last[21], 4),
            "host_completion_order": round(last[22], 4), "host_hierarchy_metadata": round(last[23], 4),
        },
        "counters": {
            "chunks": int(last[16]), "nodes": int(last[8]), "sampled_nodes": int(last[9]),
            "candidates": int(last[10]), "fp32_pair_tests": int(last[12]), "fp64_rechecks": int(last[11]),
            "max_fp32_rel_err": last[13], "libstdcxx_order_nodes": int(last[14]), "sequential_nodes": int(last[15]),
            "octree_rounds": int(last[17]), "depth": int(last[18]), "octree_bytes": int(last[7]),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "chunking pass (k_count + k_tile_hist + scan + k_distribute)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": chunk_bytes,
            "kernel_ms": round(kern_s * 1e3, 4),
        },
        "sampling_roofline": {
            "bound": "hbm",
            "kernel": "node sampling (k_prepare .. k_copy_tiles, three set streams)",
            "algorithmic_bytes_per_candidate": 66,
            "candidates": int(last[10]),
            "achieved": round(66 * last[10] / last[5] / 1e9, 1) if last[5] > 0 else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(66 * last[10] / last[5] / 1e9 / HBM_PEAK_GBS, 4) if last[5] > 0 else None,
        },
        "precision": precision,
        "autotune": tuned,
        "io_s": io,
        "plain_vs_batched": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    L.pcg_device_free(d_las)
    L.pcg_device_free(d_q)


if __name__ == "__main__":
    main()
