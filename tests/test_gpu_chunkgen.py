"""Phase 2 alone on pre-chunked sets of the batching workload, through the C ABI.

potreeconverter_amd.ChunkGen generates level-4 chunks "rABCD" with
log-normal point counts and three extra-bytes attributes (34-byte records),
either as a chunks/ directory (host) or straight into HBM (GPU).  The
fixtures c2_*.json were written by the reference indexer on the host-written
directories (tests/golden/make_golden.py); the GPU indexer must reproduce
them from the directory (pcg_do_indexing) and from device-resident records
(pcg_do_indexing_device), and match the oracle on the full 4096-chunk shape
at a reduced point count.
"""
import ctypes
import filecmp
import glob
import json
import os

import pytest

import potreeconverter_amd as pc
from helpers import digest_output, oracle_index

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "c2_*.json")))
IDS = [os.path.basename(p)[:-5] for p in FIXTURES]


def _load(path):
    with open(path) as f:
        return json.load(f)


def _check_against(fx, out):
    with open(os.path.join(out, "metadata.json")) as f:
        assert f.read() == fx["metadata_json"]
    got = digest_output(str(out))
    assert set(got) == set(fx["nodes"])
    bad = [n for n in got if got[n] != fx["nodes"][n]]
    assert not bad, f"{len(bad)} nodes differ, e.g. {bad[:5]}"


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        assert pc.lib().pcg_device_alloc(ctypes.byref(self.ptr), nbytes + 64) == 0
        self.nbytes = nbytes

    def download(self):
        out = ctypes.create_string_buffer(max(self.nbytes, 1))
        assert pc.lib().pcg_copy(out, self.ptr, self.nbytes, 1) == 0
        return out.raw[:self.nbytes]

    def free(self):
        pc.lib().pcg_device_free(self.ptr)


@pytest.mark.parametrize("fixture", FIXTURES, ids=IDS)
def test_chunks_directory_matches_reference_fixture(fixture, tmp_path):
    fx = _load(fixture)
    case = fx["case"]
    name = os.path.basename(fixture)[:-5]
    print(f"case {name} {case}", flush=True)
    g = pc.ChunkGen(**case["chunkgen"])
    g.write(tmp_path)
    pc.do_indexing(tmp_path, case["method"], name, g.total)
    _check_against(fx, tmp_path)


@pytest.mark.parametrize("fixture", FIXTURES, ids=IDS)
def test_device_chunks_match_reference_fixture(fixture, tmp_path):
    fx = _load(fixture)
    case = fx["case"]
    name = os.path.basename(fixture)[:-5]
    print(f"device case {name} {case}", flush=True)
    g = pc.ChunkGen(**case["chunkgen"])
    buf = DeviceBuffer(g.total * g.RECORD_BYTES)
    try:
        meta = g.device(buf.ptr.value)
        # the GPU generator writes the host generator's records and metadata
        g.write(tmp_path / "host")
        with open(tmp_path / "host" / "chunks" / "metadata.json") as f:
            assert f.read() == meta
        host = b"".join(open(tmp_path / "host" / "chunks" / (n + ".bin"), "rb").read() for n in g.names)
        assert buf.download() == host
        out = tmp_path / "out"
        st = pc.do_indexing_device(buf.ptr.value, g.names, g.counts, meta, case["method"], name=name, outdir=out)
        _check_against(fx, out)
        assert st[16] == g.num_chunks
        # the input records are only read
        assert buf.download() == host
    finally:
        buf.free()


@pytest.mark.parametrize("method", ["poisson_average", "poisson"])
def test_full_chunk_grid_matches_oracle(method, tmp_path):
    """All 4096 level-4 chunks (the batching workload's shape) at a reduced
    total: median 600 points, clipped to 100 .. 200 000."""
    g = pc.ChunkGen(4096, 600, 1.16, 100, 200_000, seed=11)
    print(f"full grid {method}: {g.total} points", flush=True)
    cpu = tmp_path / "cpu"
    g.write(cpu)
    oracle_index(cpu, method, "c2")
    buf = DeviceBuffer(g.total * g.RECORD_BYTES)
    try:
        meta = g.device(buf.ptr.value)
        gpu = tmp_path / "gpu"
        pc.do_indexing_device(buf.ptr.value, g.names, g.counts, meta, method, name="c2", outdir=gpu)
        # a second call on the same input gives the same files
        gpu2 = tmp_path / "gpu2"
        pc.do_indexing_device(buf.ptr.value, g.names, g.counts, meta, method, name="c2", outdir=gpu2)
    finally:
        buf.free()
    for f in ("octree.bin", "hierarchy.bin", "metadata.json"):
        assert filecmp.cmp(cpu / f, gpu / f, shallow=False), f
        assert filecmp.cmp(gpu / f, gpu2 / f, shallow=False), f


def test_sample_sets_on_the_full_chunk_grid(tmp_path):
    """4096 chunk subtrees in 64 sets on 3 streams and in 7 sets on one
    stream: the oracle's bytes (Poisson-average)."""
    g = pc.ChunkGen(4096, 300, 1.16, 50, 100_000, seed=12)
    cpu = tmp_path / "cpu"
    g.write(cpu)
    oracle_index(cpu, "poisson_average", "c2")
    buf = DeviceBuffer(g.total * g.RECORD_BYTES)
    L = pc.lib()
    try:
        meta = g.device(buf.ptr.value)
        for streams, sets in ((3, 64), (1, 7)):
            print(f"full grid sets: streams={streams} sets={sets}", flush=True)
            assert L.pcg_set_option(b"sample_groups", streams) == 0
            assert L.pcg_set_option(b"sample_sets", sets) == 0
            out = tmp_path / f"s{streams}_{sets}"
            pc.do_indexing_device(buf.ptr.value, g.names, g.counts, meta, "poisson_average", name="c2", outdir=out)
            for f in ("octree.bin", "hierarchy.bin", "metadata.json"):
                assert filecmp.cmp(cpu / f, out / f, shallow=False), (streams, sets, f)
    finally:
        L.pcg_set_option(b"sample_groups", 0)
        L.pcg_set_option(b"sample_sets", 0)
        buf.free()


def test_device_indexing_rejects_bad_chunk_lists(tmp_path):
    g = pc.ChunkGen(8, 500, 0.5, 100, 1000, seed=1)
    buf = DeviceBuffer(g.total * g.RECORD_BYTES)
    try:
        meta = g.device(buf.ptr.value)
        with pytest.raises(pc.PotreeGpuError, match="ascending"):
            pc.do_indexing_device(buf.ptr.value, g.names[::-1], g.counts[::-1], meta, "poisson")
        with pytest.raises(pc.PotreeGpuError, match="bad chunk"):
            pc.do_indexing_device(buf.ptr.value, ["x0000"] + g.names[1:], g.counts, meta, "poisson")
    finally:
        buf.free()
