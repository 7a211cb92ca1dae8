"""The synthetic LAS generator: deterministic, bounds-exact headers."""
import struct

import numpy as np

import potreeconverter_amd as pc
from helpers import read


def test_deterministic_and_seeded():
    a = pc.LasGen(0, 5000, seed=1).records()
    b = pc.LasGen(0, 5000, seed=1).records()
    c = pc.LasGen(0, 5000, seed=2).records()
    assert a == b and a != c
    # a prefix of a longer stream with the same tiling is the same stream
    g = pc.LasGen(0, 10000, seed=1)
    assert g.records(0, 3000) + g.records(3000, 2000) == g.records(0, 5000)


def test_header_bounds_are_exact(tmp_path):
    g = pc.LasGen(1, 20000, seed=3)
    p = tmp_path / "g.las"
    g.write(p)
    raw = read(p)
    n = struct.unpack_from("<Q", raw, 247)[0]
    scale = np.array(struct.unpack_from("<3d", raw, 131))
    off = np.array(struct.unpack_from("<3d", raw, 155))
    mx = np.array([struct.unpack_from("<d", raw, o)[0] for o in (179, 195, 211)])
    mn = np.array([struct.unpack_from("<d", raw, o)[0] for o in (187, 203, 219)])
    recs = np.frombuffer(raw[375:375 + n * 26], dtype=np.uint8).reshape(n, 26)
    xyz = recs[:, :12].copy().view(np.int32).reshape(n, 3)
    x = scale * xyz + off
    assert np.array_equal(x.min(axis=0), mn) and np.array_equal(x.max(axis=0), mx)
    assert mx[0] - mn[0] > 4000 and mn[0] >= 500000  # 5 km tile in UTM-like coordinates
    cls = recs[:, 15]
    assert set(np.unique(cls)) <= {2, 6} and (cls == 6).any()


def test_precision_scene():
    """Preset 2: anchors fix the X/Y range to 8 192 000 units of 0.0001 m,
    about 10% of the points share a quantized position up to +-1 unit with
    another, and 5% sit on a lattice of pitch 4000 units (0.4 m, the node
    spacing at level 4 of a 819.2 m cube)."""
    n = 200_000
    g = pc.LasGen(2, n, seed=3)
    recs = np.frombuffer(g.records(), dtype=np.uint8).reshape(n, 26)
    xyz = recs[:, :12].copy().view(np.int32).reshape(n, 3)
    assert xyz[:, 0].min() == 0 and xyz[:, 0].max() == 8_192_000
    assert xyz[:, 1].min() == 0 and xyz[:, 1].max() == 8_192_000
    assert tuple(xyz[0, :2]) == (0, 0) and tuple(xyz[1, :2]) == (8_192_000, 8_192_000)
    x0 = 8_192_000 // 4
    on = ((xyz[:, 0] - x0) % 4000 == 0) & ((xyz[:, 1] - x0) % 4000 == 0) & (xyz[:, 0] >= x0) & (xyz[:, 1] >= x0)
    on[:2] = False  # the anchors
    assert 0.045 < on.mean() < 0.055
    lat = xyz[on]
    assert len(np.unique(lat[:, 2])) == 1  # one flat plane
    # near-duplicates: a position within +-1 unit of another point's
    rest = xyz[~on].astype(np.int64)
    _, counts = np.unique(rest // 4, axis=0, return_counts=True)
    assert counts[counts > 1].sum() > 0.07 * n
    # spacing at level 4 of the cube equals the lattice pitch up to binary64 rounding
    cube = (8_192_000 * 1e-4 + 1e6) - (0 * 1e-4 + 1e6)
    sp = cube / 128 / 2**4
    assert abs(sp - 0.4) < 1e-12 and sp != 0.4


C2_SIGMA = 1.16  # median 50 000, clip 1 000 .. 5 000 000: about 400M points over 4096 chunks


def test_chunkgen_counts_and_names():
    g = pc.ChunkGen(4096, 50_000, C2_SIGMA, 1_000, 5_000_000)
    assert g.names == sorted(g.names) and len(set(g.names)) == 4096
    assert g.names[0] == "r0000" and g.names[-1] == "r7777"
    c = np.array(g.counts)
    assert c.min() >= 1_000 and c.max() <= 5_000_000
    assert abs(np.median(c) / 50_000 - 1) < 0.06
    # log(count) ~ N(log 50000, sigma) away from the clip
    s = np.log(c[(c > 1_000) & (c < 5_000_000)]).std()
    assert abs(s / C2_SIGMA - 1) < 0.06
    assert 3.5e8 < g.total < 4.5e8
    # a subset of chunks spreads over the cube: stride 4096 / num_chunks cells
    h = pc.ChunkGen(64, 1000, 0.5, 10, 5000)
    assert h.names[1] == "r0100" and h.names[8] == "r1000"


def test_chunkgen_records_stay_in_their_cell(tmp_path):
    g = pc.ChunkGen(16, 3000, 0.5, 100, 20_000, seed=5)
    g.write(tmp_path)
    w = 2048.0 / 16
    for i in (0, 5, 15):
        raw = open(tmp_path / "chunks" / (g.names[i] + ".bin"), "rb").read()
        assert raw == g.records(i)
        assert g.records(i, 10, 7) == raw[10 * 34:17 * 34]
        rec = np.frombuffer(raw, np.uint8).reshape(-1, 34)
        xyz = rec[:, :12].copy().view(np.int32) * 0.001
        # the cell from the name digits (x -> 4, y -> 2, z -> 1 per level)
        cell = [0, 0, 0]
        for d in g.names[i][1:]:
            d = int(d)
            cell = [2 * cell[0] + (d >> 2), 2 * cell[1] + ((d >> 1) & 1), 2 * cell[2] + (d & 1)]
        lo = np.array(cell) * w
        assert (xyz >= lo).all() and (xyz < lo + w).all()
        assert set(np.unique(rec[:, 16])) <= {2, 6}
        gauss = rec[:, 27:31].copy().view(np.float32).ravel()
        assert abs(gauss.mean()) < 0.1 and abs(gauss.std() - 1) < 0.1
    meta = open(tmp_path / "chunks" / "metadata.json").read()
    assert '"name": "gauss"' in meta and '"name": "uniform8"' in meta


def test_serving_job_stream():
    """The serving loop's job stream: Poisson arrivals at the given rate and
    log-uniform sizes, reproducible from the seed."""
    from potreeconverter_amd.serve import make_jobs, percentile
    a = make_jobs(4000, 20.0, 100_000, 50_000_000, 3)
    b = make_jobs(4000, 20.0, 100_000, 50_000_000, 3)
    assert [(j.arrival, j.points) for j in a] == [(j.arrival, j.points) for j in b]
    gaps = np.diff([0.0] + [j.arrival for j in a])
    assert abs(gaps.mean() * 20.0 - 1) < 0.05 and abs(gaps.std() / gaps.mean() - 1) < 0.06
    logn = np.log([j.points for j in a])
    assert logn.min() >= np.log(100_000) - 1e-9 and logn.max() <= np.log(50_000_000) + 1e-9
    # log-uniform: each decade-ish quarter of the log range holds a quarter
    q = np.histogram(logn, bins=4, range=(np.log(100_000), np.log(50_000_000)))[0] / len(a)
    assert np.all(np.abs(q - 0.25) < 0.03)
    assert percentile([1, 2, 3, 4], 50) == 2.5 and percentile([5], 99) == 5
