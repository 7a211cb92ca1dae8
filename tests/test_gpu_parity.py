"""gfx950 path against the CPU oracle, byte for byte, through the C ABI.

The oracle is pinned to the reference indexer by tests/test_oracle_golden.py;
here the GPU conversion must produce the same octree.bin, hierarchy.bin and
metadata.json bytes (both sides write the canonical node order), and the GPU
chunker the same chunk files.
"""
import ctypes
import filecmp
import json
import os

import numpy as np
import pytest

import potreeconverter_amd as pc
from helpers import digest_output, make_las, oracle, oracle_convert, read
from scenes import make_scene

pytestmark = pytest.mark.gpu

CASES = [
    # preset, points, seed, extra bytes
    (0, 150_000, 2, False),
    (0, 2_000_000, 0, False),
    (0, 2_000_000, 7, False),
    (0, 3_000_000, 1, False),
    (0, 400_000, 5, True),
    (1, 1_000_000, 1, False),
    (2, 1_000_000, 3, False),
    (2, 300_000, 11, True),
]



def same_files(a, b, names):
    return {n: filecmp.cmp(os.path.join(a, n), os.path.join(b, n), shallow=False) for n in names}


@pytest.mark.parametrize("preset,n,seed,eb", CASES, ids=[f"p{c[0]}_{c[1]}_s{c[2]}{'_eb' if c[3] else ''}" for c in CASES])
@pytest.mark.parametrize("method", ["poisson", "poisson_average"])
def test_convert_matches_oracle(preset, n, seed, eb, method, tmp_path):
    print(f"case preset={preset} n={n} seed={seed} extra={eb} method={method}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, preset, n, seed, eb)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, method)
    pc.convert(las, gpu, method)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res


FORMAT_CASES = [(f, eb) for f in range(8) for eb in (False, True)]


@pytest.mark.parametrize("fmt,eb", FORMAT_CASES, ids=[f"pdrf{f}{'_eb' if e else ''}" for f, e in FORMAT_CASES])
def test_chunking_matches_oracle(fmt, eb, tmp_path):
    """Chunk files and chunks/metadata.json of every point format 0-7, with
    and without extra bytes (the statistics, transcode and classification
    histogram of each format's attribute handlers)."""
    print(f"chunking case pdrf={fmt} extra={eb}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 1, 800_000, 4 + fmt, eb, fmt)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, "poisson", chunk_only=True)
    pc.do_chunking(las, gpu)
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b
    assert all(same_files(cpu / "chunks", gpu / "chunks", a).values())


SHUFFLED_CASES = [(2, False), (7, True), (5, False)]


@pytest.mark.parametrize("fmt,eb", SHUFFLED_CASES, ids=[f"pdrf{f}{'_eb' if e else ''}" for f, e in SHUFFLED_CASES])
def test_chunking_of_shuffled_records_matches_oracle(fmt, eb, tmp_path):
    """Records in random order, so that every 64-point step of the scatter
    holds points of many chunks (the per-step run matching past its first
    rounds); formats 2 and 7 take the dword scatter, format 5 the byte-wise
    one."""
    print(f"shuffled chunking case pdrf={fmt} extra={eb}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 1, 400_000, 40 + fmt, eb, fmt)
    data = bytearray(las.read_bytes())
    assert data[24:26] == bytes([1, 4])
    off = int.from_bytes(data[96:100], "little")
    rl = int.from_bytes(data[105:107], "little")
    n = int.from_bytes(data[247:255], "little")
    assert n == 400_000 and off + n * rl <= len(data)
    recs = np.frombuffer(bytes(data[off:off + n * rl]), np.uint8).reshape(n, rl)
    data[off:off + n * rl] = recs[np.random.default_rng(fmt).permutation(n)].tobytes()
    las.write_bytes(bytes(data))
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, "poisson", chunk_only=True)
    pc.do_chunking(las, gpu)
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b and len(a) > 8
    assert all(same_files(cpu / "chunks", gpu / "chunks", a).values())


FORMAT_CONVERT_CASES = [(1, "poisson", False), (3, "poisson_average", True), (4, "poisson", False),
                        (5, "poisson_average", False), (6, "poisson", True), (7, "poisson_average", False)]


@pytest.mark.parametrize("fmt,method,eb", FORMAT_CONVERT_CASES,
                         ids=[f"pdrf{c[0]}_{c[1]}{'_eb' if c[2] else ''}" for c in FORMAT_CONVERT_CASES])
def test_convert_point_formats_matches_oracle(fmt, method, eb, tmp_path):
    print(f"format case pdrf={fmt} method={method} extra={eb}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 0, 600_000, 20 + fmt, eb, fmt)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, method)
    pc.convert(las, gpu, method)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res


def test_indexing_of_reference_format_chunks(tmp_path):
    """doIndexing on chunks/ written to the reference's file contract."""
    las = tmp_path / "in.las"
    make_las(las, 0, 900_000, 8)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "poisson", keep_chunks=True)
    gpu = tmp_path / "gpu"
    os.makedirs(gpu)
    os.system(f"cp -r {cpu}/chunks {gpu}/")
    pc.do_indexing(gpu, "poisson", "in", 900_000)
    assert all(same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"]).values())


def test_count_cells_entry_point():
    L = pc.lib()
    g = pc.LasGen(0, 300_000, seed=6)
    recs = g.records()
    n, rl = 300_000, 26
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), len(recs) + 64) == 0
    assert L.pcg_copy(d, recs, len(recs), 0) == 0
    a = np.frombuffer(recs, dtype=np.uint8).reshape(n, rl)[:, :12].copy().view(np.int32).reshape(n, 3)
    scale, off = [0.001] * 3, [0.0] * 3
    mn = [(a[:, k].min() * 0.001) for k in range(3)]
    mx = [(a[:, k].max() * 0.001) for k in range(3)]
    dp = lambda v: (ctypes.c_double * 3)(*v)
    G = 128
    dg = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(dg), 4 * G ** 3) == 0
    rc = L.pcg_count_cells(d, n, rl, dp(scale), dp(off), dp(scale), dp(mn), dp(mn), dp(mx), G, dg, None)
    assert rc == 0, L.pcg_last_error()
    got = np.zeros(G ** 3, dtype=np.int32)
    assert L.pcg_copy(got.ctypes.data_as(ctypes.c_void_p), dg, 4 * G ** 3, 1) == 0
    want = np.zeros(G ** 3, dtype=np.int32)
    buf = ctypes.create_string_buffer(recs, len(recs))
    rc = oracle().orc_count_cells(buf, n, rl, dp(scale), dp(off), dp(scale), dp(mn), dp(mn), dp(mx), G,
                                  want.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    assert np.array_equal(got, want)
    L.pcg_device_free(d)
    L.pcg_device_free(dg)


def test_out_of_box_point_is_rejected_like_the_reference(tmp_path):
    """A header whose bounds exclude a point fails with the reference's 123."""
    las = tmp_path / "bad.las"
    make_las(las, 0, 50_000, 1)
    raw = bytearray(read(las))
    # shrink max x below the data
    import struct
    mx = struct.unpack_from("<d", raw, 179)[0]
    mnx = struct.unpack_from("<d", raw, 187)[0]
    struct.pack_into("<d", raw, 179, mnx + (mx - mnx) * 0.5)
    # keep the cube from the other axes small too
    for o_max, o_min in ((195, 203), (211, 219)):
        lo = struct.unpack_from("<d", raw, o_min)[0]
        struct.pack_into("<d", raw, o_max, lo + (mx - mnx) * 0.25)
    las.write_bytes(bytes(raw))
    t = (ctypes.c_double * 8)()
    rc = pc.lib().pcg_do_chunking(str(las).encode(), str(tmp_path / "o").encode(), t)
    assert rc == 123


@pytest.mark.parametrize("preset", [1, 2])
def test_gpu_generator_equals_cpu_generator(preset):
    L = pc.lib()
    g = pc.LasGen(preset, 1_000_000, seed=7)
    cpu = g.records()
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), len(cpu) + 64) == 0
    q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    dq = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(dq), 64) == 0
    assert L.pcg_copy(dq, q, 24, 0) == 0
    assert L.pcg_lasgen_device(g.buf, 0, 1_000_000, d, dq) == 0
    out = ctypes.create_string_buffer(len(cpu))
    assert L.pcg_copy(out, d, len(cpu), 1) == 0
    assert out.raw == cpu
    L.pcg_device_free(d)
    L.pcg_device_free(dq)


def test_large_run_is_deterministic_and_conserves_points(tmp_path):
    """20M points: two device runs give identical bytes; Poisson keeps every point once."""
    L = pc.lib()
    n = 20_000_000
    g = pc.LasGen(1, n, seed=0)
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), n * 26 + 64) == 0
    q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    dq = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(dq), 64) == 0
    assert L.pcg_copy(dq, q, 24, 0) == 0
    assert L.pcg_lasgen_device(g.buf, 0, n, d, dq) == 0
    assert L.pcg_copy(q, dq, 24, 1) == 0
    hdr = ctypes.create_string_buffer(375)
    assert L.pcg_lasgen_header(g.buf, q, hdr) == 0
    stats = (ctypes.c_double * 24)()
    outs = []
    for k in range(2):
        o = tmp_path / f"run{k}"
        rc = L.pcg_convert_device(d, n, hdr, b"poisson", str(o).encode(), stats)
        assert rc == 0, L.pcg_last_error()
        outs.append(o)
    assert all(same_files(outs[0], outs[1], ["octree.bin", "hierarchy.bin", "metadata.json"]).values())
    meta = json.load(open(outs[0] / "metadata.json"))
    assert meta["points"] == n
    assert os.path.getsize(outs[0] / "octree.bin") == 27 * n
    L.pcg_device_free(d)
    L.pcg_device_free(dq)


def _set(key, value):
    rc = pc.lib().pcg_set_option(key.encode(), value)
    assert rc == 0, pc.lib().pcg_last_error()


RING_CASES = [(0, 150_000, 2, False), (0, 2_000_000, 7, False), (0, 400_000, 5, True), (1, 1_000_000, 1, False),
              (2, 1_000_000, 3, False)]


@pytest.mark.parametrize("preset,n,seed,eb", RING_CASES, ids=[f"p{c[0]}_{c[1]}_s{c[2]}" for c in RING_CASES])
def test_ring_path_matches_oracle(preset, n, seed, eb, tmp_path):
    """Every Poisson level through the accepted-shell ring kernel."""
    print(f"ring case preset={preset} n={n} seed={seed} extra={eb}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, preset, n, seed, eb)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, "poisson")
    _set("ring_min_nodes", 1)
    try:
        pc.convert(las, gpu, "poisson")
    finally:
        _set("ring_min_nodes", 512)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res


@pytest.mark.parametrize("method", ["poisson", "poisson_average"])
def test_plain_sampling_matches_oracle(method, tmp_path):
    """One launch sequence per node gives the batched path's bytes."""
    print(f"plain case method={method}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 0, 300_000, 4, False)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, method)
    _set("plain_sampling", 1)
    try:
        pc.convert(las, gpu, method)
    finally:
        _set("plain_sampling", 0)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res


def test_ring_and_cell_paths_agree_at_20m(tmp_path):
    """20M points, where the widest levels use the ring and the bucket sort
    by default: the same bytes as with every level on the cell-table path and
    bucket-sorted, and every level on the ring and radix-sorted."""
    L = pc.lib()
    n = 20_000_000
    g = pc.LasGen(1, n, seed=3)
    d = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(d), n * 26 + 64) == 0
    q = (ctypes.c_int32 * 6)(2**31 - 1, 2**31 - 1, 2**31 - 1, -2**31, -2**31, -2**31)
    dq = ctypes.c_void_p()
    assert L.pcg_device_alloc(ctypes.byref(dq), 64) == 0
    assert L.pcg_copy(dq, q, 24, 0) == 0
    assert L.pcg_lasgen_device(g.buf, 0, n, d, dq) == 0
    assert L.pcg_copy(q, dq, 24, 1) == 0
    hdr = ctypes.create_string_buffer(375)
    assert L.pcg_lasgen_header(g.buf, q, hdr) == 0
    stats = (ctypes.c_double * 24)()
    outs = []
    try:
        for k, (ring_min, bucket_min) in enumerate(((512, 512), (1 << 30, 1), (1, 1 << 30))):
            print(f"20M run ring_min_nodes={ring_min} bucket_min_nodes={bucket_min}", flush=True)
            _set("ring_min_nodes", ring_min)
            _set("bucket_min_nodes", bucket_min)
            o = tmp_path / f"run{k}"
            rc = L.pcg_convert_device(d, n, hdr, b"poisson", str(o).encode(), stats)
            assert rc == 0, L.pcg_last_error()
            outs.append(o)
    finally:
        _set("ring_min_nodes", 512)
        _set("bucket_min_nodes", 512)
        L.pcg_device_free(d)
        L.pcg_device_free(dq)
    for o in outs[1:]:
        assert all(same_files(outs[0], o, ["octree.bin", "hierarchy.bin", "metadata.json"]).values())


def test_sample_groups_agree(tmp_path):
    """Chunk subtrees sampled in 1, 3 and 8 concurrent groups: the same bytes
    (and the oracle's)."""
    las = tmp_path / "in.las"
    make_las(las, 0, 2_000_000, 7, False)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "poisson_average")
    try:
        for g in (1, 3, 8):
            print(f"groups case {g}", flush=True)
            _set("sample_groups", g)
            out = tmp_path / f"g{g}"
            pc.convert(las, out, "poisson_average")
            res = same_files(cpu, out, ["octree.bin", "hierarchy.bin", "metadata.json"])
            assert all(res.values()), (g, res)
    finally:
        _set("sample_groups", 0)


def test_sample_sets_agree(tmp_path):
    """More chunk-subtree sets than streams (sets interleaved level by level
    on a stream, sharing its buffers): the oracle's bytes."""
    las = tmp_path / "in.las"
    make_las(las, 0, 2_000_000, 9, False)
    cpu = tmp_path / "cpu"
    oracle_convert(las, cpu, "poisson")
    try:
        for groups, sets in ((3, 12), (2, 64), (1, 5)):
            print(f"sets case streams={groups} sets={sets}", flush=True)
            _set("sample_groups", groups)
            _set("sample_sets", sets)
            out = tmp_path / f"s{groups}_{sets}"
            pc.convert(las, out, "poisson")
            res = same_files(cpu, out, ["octree.bin", "hierarchy.bin", "metadata.json"])
            assert all(res.values()), (groups, sets, res)
    finally:
        _set("sample_groups", 0)
        _set("sample_sets", 0)


SCENE_CASES = [(sc, m) for sc in ("stacks", "dups", "spacing", "sphere") for m in ("poisson", "poisson_average")]


@pytest.mark.parametrize("scene,method", SCENE_CASES, ids=[f"{a}_{b}" for a, b in SCENE_CASES])
def test_edge_scene_matches_oracle(scene, method, tmp_path):
    """The edge-case scenes (pinned to the reference indexer by fixtures):
    equal-distance runs longer than a wavefront, the duplicate path of
    buildHierarchy, neighbours exactly one spacing apart, and the 10 001-check
    limit of the Poisson scan, which must route root candidates through the
    capped acceptance."""
    print(f"scene case {scene} method={method}", flush=True)
    las = tmp_path / "in.las"
    n = make_scene(scene, las)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, method)
    t = pc.convert(las, gpu, method)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res
    if scene == "stacks":
        assert t[6] > 0  # nodes put in libstdc++ order
    if scene == "sphere" and method == "poisson":
        assert t[7] > 0  # nodes where the 10 001-check limit can bind
    if scene == "dups" and method == "poisson":
        # 14 999 + 10 000 duplicate records are dropped as the reference drops them
        assert os.path.getsize(gpu / "octree.bin") == 27 * (n - 24_999)


GRID_CASES = [(256, 100_000), (512, 100_000), (256, 1_000_000)]


@pytest.mark.parametrize("grid,cap", GRID_CASES, ids=[f"g{g}_cap{c}" for g, c in GRID_CASES])
def test_forced_chunk_grid_matches_oracle(grid, cap, tmp_path):
    """The chunk grids the reference picks for 100M-500M points (256) and
    above (512), forced on 2M points of the bench's generator in both the
    oracle and the library: chunk files, chunks/metadata.json and the octree
    bytes.  (A cap at or above the point count is no configuration of the
    reference: its LUT then holds no chunk.)"""
    print(f"forced grid case G={grid} cap={cap}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 1, 2_000_000, 0, False)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle().orc_set_chunking(grid, cap)
    try:
        oracle_convert(las, cpu, "poisson", keep_chunks=True)
    finally:
        oracle().orc_set_chunking(0, 0)
    _set("chunk_grid", grid)
    _set("chunk_cap", cap)
    try:
        pc.convert(las, gpu, "poisson", keep_chunks=True)
    finally:
        _set("chunk_grid", 0)
        _set("chunk_cap", 0)
    res = same_files(cpu, gpu, ["octree.bin", "hierarchy.bin", "metadata.json"])
    assert all(res.values()), res
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b
    assert all(same_files(cpu / "chunks", gpu / "chunks", a).values())


@pytest.mark.parametrize("fmt,eb", [(2, False), (7, True)], ids=["pdrf2", "pdrf7_eb"])
def test_bytewise_distribute_matches_oracle(fmt, eb, tmp_path):
    """The byte-wise scatter (taken by point formats 4-5, records over 64
    bytes and very many chunks) forced on the layouts the tile transcode
    normally takes: the same chunk files."""
    print(f"byte-wise distribute case pdrf={fmt} extra={eb}", flush=True)
    las = tmp_path / "in.las"
    make_las(las, 1, 500_000, 60 + fmt, eb, fmt)
    cpu, gpu = tmp_path / "cpu", tmp_path / "gpu"
    oracle_convert(las, cpu, "poisson", chunk_only=True)
    _set("distribute_tail", 0)
    try:
        pc.do_chunking(las, gpu)
    finally:
        _set("distribute_tail", 1)
    a, b = sorted(os.listdir(cpu / "chunks")), sorted(os.listdir(gpu / "chunks"))
    assert a == b and len(a) > 4
    assert all(same_files(cpu / "chunks", gpu / "chunks", a).values())
