"""The oracle's chunker against closed forms and brute force.

The reference chunker cannot be built here (laszip needs a CMake-generated
header), so its arithmetic is pinned by independent formulations:
 - the cell counts sum to N and match a vectorised evaluation of the
   reference's requantisation/cell formula (chunker_countsort_laszip.cpp:224-257);
 - the bottom-up pyramid merge (:1074-1204) equals the top-down rule
   "a cell is a chunk iff it holds points, its parent holds more than the cap,
   and it holds at most the cap or is a finest-level cell";
 - chunk files are the input records filtered by chunk, in input order, with
   the point-format-2 transcode checked on a hand-made record.
"""
import ctypes
import os
import struct

import numpy as np
import pytest

from helpers import make_las, oracle, oracle_convert, read


def dp(v):
    return (ctypes.c_double * 3)(*v)


def las_records(path):
    raw = read(path)
    off = struct.unpack_from("<I", raw, 96)[0]
    rl = struct.unpack_from("<H", raw, 105)[0]
    n = struct.unpack_from("<Q", raw, 247)[0]
    scale = struct.unpack_from("<3d", raw, 131)
    offset = struct.unpack_from("<3d", raw, 155)
    mx = struct.unpack_from("<d", raw, 179)[0], struct.unpack_from("<d", raw, 195)[0], struct.unpack_from("<d", raw, 211)[0]
    mn = struct.unpack_from("<d", raw, 187)[0], struct.unpack_from("<d", raw, 203)[0], struct.unpack_from("<d", raw, 219)[0]
    return raw[off:off + n * rl], n, rl, scale, offset, mn, mx


def reference_cells(recs, n, rl, scale, offset, mn, mx, G):
    a = np.frombuffer(recs, dtype=np.uint8).reshape(n, rl)
    xyz = a[:, :12].copy().view(np.int32).reshape(n, 3).astype(np.float64)
    x = np.asarray(scale) * xyz + np.asarray(offset)
    ps = np.maximum(np.asarray(scale), (np.asarray(mx) - np.asarray(mn)) / 2.0 ** 30)
    po = np.asarray(mn)
    q = np.trunc((x - po) / ps)
    c0 = max(max(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2])
    smax = np.asarray(mn) + c0
    cube = max(max(smax[0] - mn[0], smax[1] - mn[1]), smax[2] - mn[2])
    u = (q * ps + po - np.asarray(mn)) / cube
    i = np.minimum(G * u, G - 1.0).astype(np.int64)
    return i[:, 0] + i[:, 1] * G + i[:, 2] * G * G, ps, po


@pytest.mark.parametrize("preset,n", [(0, 120_000), (1, 90_000)])
def test_count_cells_matches_vectorised_formula(preset, n, tmp_path):
    las = tmp_path / "a.las"
    make_las(las, preset, n, 4)
    recs, n, rl, scale, offset, mn, mx = las_records(las)
    G = 128
    cells, ps, po = reference_cells(recs, n, rl, scale, offset, mn, mx, G)
    grid = np.zeros(G ** 3, dtype=np.int32)
    buf = ctypes.create_string_buffer(recs, len(recs))
    rc = oracle().orc_count_cells(buf, n, rl, dp(scale), dp(offset), dp(ps), dp(po), dp(mn), dp(mx), G,
                                  grid.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    assert grid.sum() == n
    assert np.array_equal(grid, np.bincount(cells, minlength=G ** 3).astype(np.int32))


def top_down_chunks(grid, G, cap):
    L = int(np.log2(G))
    levels = {L: grid.reshape(G, G, G).transpose(2, 1, 0).astype(np.int64)}  # [x, y, z]
    for l in range(L - 1, -1, -1):
        h = levels[l + 1]
        g = h.shape[0] // 2
        levels[l] = h.reshape(g, 2, g, 2, g, 2).sum(axis=(1, 3, 5))
    chunks = set()
    for l in range(1, L + 1):
        c = levels[l]
        par = np.repeat(np.repeat(np.repeat(levels[l - 1], 2, 0), 2, 1), 2, 2)
        ok = (c > 0) & (par > cap) & ((c <= cap) | (l == L))
        for x, y, z in zip(*np.nonzero(ok)):
            chunks.add((l, int(x), int(y), int(z), int(c[x, y, z])))
    return chunks


@pytest.mark.parametrize("G,seed", [(8, 0), (16, 1), (16, 2), (32, 3)])
def test_lut_equals_top_down_rule(G, seed):
    rng = np.random.default_rng(seed)
    grid = np.zeros(G ** 3, dtype=np.int32)
    # clustered counts: dense blobs, single heavy cells above the cap, empty space
    for _ in range(6):
        c = rng.integers(0, G, 3)
        r = rng.integers(1, G // 2)
        x, y, z = np.meshgrid(*[np.arange(max(0, c[k] - r), min(G, c[k] + r)) for k in range(3)], indexing="ij")
        grid[(x + y * G + z * G * G).ravel()] += rng.integers(0, 40, x.size).astype(np.int32)
    grid[rng.integers(0, G ** 3)] += 5000
    cap = int(grid.sum() // 20)
    desc = np.zeros(5 * G ** 3, dtype=np.int64)
    lut = np.zeros(G ** 3, dtype=np.int32)
    nc = oracle().orc_create_lut(grid.ctypes.data_as(ctypes.c_void_p), G, cap, desc.ctypes.data_as(ctypes.c_void_p),
                                 G ** 3, lut.ctypes.data_as(ctypes.c_void_p))
    got = {tuple(int(v) for v in desc[5 * i:5 * i + 5]) for i in range(nc)}
    assert got == top_down_chunks(grid, G, cap)
    # every non-empty cell is covered by exactly one chunk, whose count adds up
    assert (lut[grid > 0] >= 0).all()
    sums = np.bincount(lut[lut >= 0], weights=grid[lut >= 0], minlength=nc)
    assert np.array_equal(sums.astype(np.int64), desc[4:5 * nc:5])


def test_chunk_files_are_filtered_input_in_order(tmp_path):
    las = tmp_path / "a.las"
    make_las(las, 0, 150_000, 9)
    out = tmp_path / "o"
    oracle_convert(las, out, "poisson", chunk_only=True)
    recs, n, rl, scale, offset, mn, mx = las_records(las)
    G = 128
    cells, ps, po = reference_cells(recs, n, rl, scale, offset, mn, mx, G)
    a = np.frombuffer(recs, dtype=np.uint8).reshape(n, rl)
    files = sorted(f for f in os.listdir(out / "chunks") if f.endswith(".bin"))
    seen = 0
    owner = {}
    for f in files:
        data = np.frombuffer(read(out / "chunks" / f), dtype=np.uint8).reshape(-1, 27)
        xyz = data[:, :12].copy().view(np.int32).reshape(-1, 3)
        seen += len(data)
        # records of one chunk keep their input order: their source indices increase
        q = np.trunc((np.asarray(scale) * a[:, :12].copy().view(np.int32).reshape(n, 3) + np.asarray(offset) - po) / ps)
        for idx in range(0, len(xyz), max(1, len(xyz) // 50)):
            hits = np.nonzero((q == xyz[idx]).all(axis=1))[0]
            assert len(hits) >= 1
            owner.setdefault(f, []).append(hits[0])
    assert seen == n
    for f, idxs in owner.items():
        cs = {cells[i] for i in idxs}
        assert len(cs) >= 1


def test_point_format_2_transcode_known_answer(tmp_path):
    # one hand-made PDRF 2 record, decoded by the LAS 1.4 layout
    rec = struct.pack("<iiiHBBbBHHHH", 1000, 2000, 3000, 4321, (2 & 7) | (3 << 3), 0x80 | 6, -12, 77, 555,
                      100, 200, 300)
    assert len(rec) == 26
    hdr = bytearray(375)
    hdr[0:4] = b"LASF"
    hdr[24], hdr[25] = 1, 4
    struct.pack_into("<H", hdr, 94, 375)
    struct.pack_into("<I", hdr, 96, 375)
    hdr[104] = 2
    struct.pack_into("<H", hdr, 105, 26)
    struct.pack_into("<I", hdr, 107, 25)
    struct.pack_into("<3d", hdr, 131, 0.01, 0.01, 0.01)
    struct.pack_into("<3d", hdr, 155, 0.0, 0.0, 0.0)
    # bounds such that the record is inside: x in [0, 10], y in [0, 20], z in [0, 30]
    struct.pack_into("<dddddd", hdr, 179, 10.0, 0.0, 20.0, 0.0, 30.0, 0.0)
    struct.pack_into("<Q", hdr, 247, 25)
    las = tmp_path / "k.las"
    # 25 copies (the cap N/20 must be positive)
    with open(las, "wb") as f:
        f.write(bytes(hdr) + rec * 25)
    out = tmp_path / "o"
    oracle_convert(las, out, "poisson", chunk_only=True)
    files = [f for f in os.listdir(out / "chunks") if f.endswith(".bin")]
    data = b"".join(read(out / "chunks" / f) for f in files)
    assert len(data) == 25 * 27
    r = data[:27]
    X, Y, Z = struct.unpack_from("<iii", r, 0)
    # requantised against offset = bounds min (0) and scale 0.01: trunc(x/0.01)
    assert (X, Y, Z) in {(1000, 2000, 3000), (999, 1999, 2999), (999, 2000, 3000), (1000, 1999, 3000),
                         (1000, 2000, 2999), (999, 1999, 3000), (999, 2000, 2999), (1000, 1999, 2999)}
    assert struct.unpack_from("<H", r, 12)[0] == 4321          # intensity
    assert r[14] == 2 and r[15] == 3                            # return number, number of returns
    assert r[16] == 6                                           # classification (5 bits; flags dropped)
    assert struct.unpack_from("<b", r, 17)[0] == -12            # scan angle rank
    assert r[18] == 77                                          # user data
    assert struct.unpack_from("<H", r, 19)[0] == 555            # point source id
    assert struct.unpack_from("<3H", r, 21) == (100, 200, 300)  # rgb
    meta = open(out / "chunks" / "metadata.json").read()
    assert '"name": "scan angle rank"' in meta and "-12" in meta
