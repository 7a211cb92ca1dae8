"""Parity at the benchmarked configuration: 200M points of the bench's
stream (LAS generator preset 1, seed 0: a 5 x 5 km tile at UTM offsets,
scale 0.001, point format 2), converted with -m poisson by the library from
records in HBM (pcg_convert_device, the call bench.py times; the reference's
rule picks grid 256 and a 10M-point chunk cap at this size,
chunker_countsort_laszip.cpp:1209-1219).

The CPU side is the reference's own split: doChunking by the oracle on the
same records written as a LAS file, then doIndexing by the reference indexer
compiled from the reference sources (oracle/_ref/ref_indexer), which writes
its nodes in thread-completion order, so the two outputs are compared node
by node (name, type, child mask, point count, byte size and a digest of the
node's octree.bin bytes) plus metadata.json as text.

The reference's octree.bin is not deterministic: Writer::writeAndUnload
reserves a node's range of the shared write buffer under its lock and copies
the node's points into it after releasing the lock
(Converter/src/indexer.cpp:1496-1519), while the writer thread may already
be writing that buffer to the file (:1525-1546).  A node copied late keeps
the buffer's previous contents (zeros or older data) in the file; run twice
on the same chunks with 20 threads on fewer CPUs, the reference gives a few
such nodes per run, different ones each time.  A node whose bytes differ from
the library's in the first reference run is therefore compared with a second
run; its hierarchy record must match in both."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import potreeconverter_amd as pc
from helpers import REF_INDEXER, digest_output, oracle_convert, read

pytestmark = pytest.mark.gpu

N = 200_000_000


def _reference_indexer(outdir):
    subprocess.run([REF_INDEXER, str(outdir), "poisson", "device", str(N), "DEFAULT"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


@pytest.mark.timeout(900)
@pytest.mark.skipif(not os.path.exists(REF_INDEXER), reason="oracle/_ref/ref_indexer not built (oracle/build_ref.sh)")
def test_bench_configuration_matches_reference_indexer(tmp_path):
    print(f"bench configuration case: preset 1, seed 0, {N} points, poisson", flush=True)
    L = pc.lib()
    g = pc.LasGen(1, N, seed=0)
    rl = g.record_length
    d = ctypes.c_void_p()
    dq = ctypes.c_void_p()
    pc._check(L.pcg_device_alloc(ctypes.byref(d), N * rl + 64))
    pc._check(L.pcg_device_alloc(ctypes.byref(dq), 64))
    try:
        q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
        pc._check(L.pcg_copy(dq, q, 24, 0))
        pc._check(L.pcg_lasgen_device(g.buf, 0, N, d, dq))
        pc._check(L.pcg_copy(q, dq, 24, 1))
        hdr = ctypes.create_string_buffer(375)
        pc._check(L.pcg_lasgen_header(g.buf, q, hdr))
        gpu = tmp_path / "gpu"
        stats = (ctypes.c_double * 24)()
        pc._check(L.pcg_convert_device(d, N, hdr, b"poisson", str(gpu).encode(), stats))
        print("gpu conversion done", flush=True)
        recs = np.empty(N * rl, np.uint8)
        pc._check(L.pcg_copy(recs.ctypes.data, d, N * rl, 1))
    finally:
        L.pcg_device_free(d)
        L.pcg_device_free(dq)
    assert os.path.getsize(gpu / "octree.bin") == 27 * N
    las = tmp_path / "bench.las"
    with open(las, "wb") as f:
        f.write(hdr.raw)
        recs.tofile(f)
    del recs
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "poisson", keep_chunks=True, chunk_only=True)
    os.remove(las)
    print("oracle chunking done", flush=True)
    _reference_indexer(cpu)
    print("reference indexing done", flush=True)
    assert read(gpu / "metadata.json") == read(cpu / "metadata.json")
    mine, ref = digest_output(gpu), digest_output(cpu)
    assert set(mine) == set(ref)
    records = [n for n in mine if mine[n][:4] != ref[n][:4]]
    assert not records, [(n, mine[n][:4], ref[n][:4]) for n in records[:10]]
    differ = sorted(n for n in mine if mine[n][4] != ref[n][4])
    print(f"{len(mine)} nodes, {len(differ)} with other bytes in the first reference run: {differ[:10]}", flush=True)
    if differ:
        again = tmp_path / "cpu2"
        os.makedirs(again)
        os.symlink(cpu / "chunks", again / "chunks")
        _reference_indexer(again)
        ref2 = digest_output(again)
        assert set(ref2) == set(mine)
        wrong = [n for n in differ if mine[n] != ref2[n]]
        assert not wrong, wrong[:10]
        # a node hit by the race in the first run only
        assert len(differ) <= 64
