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
