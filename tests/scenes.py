"""Edge-case scenes for the parity tests.

Each scene is a LAS 1.4 file made from a generated tile (potreeconverter_amd
LasGen, point format 2, scale 0.001, offset 0) by moving some of its points;
every edit is seeded, so a scene name fixes the file byte for byte.  The
header bounds are recomputed from the integer positions with the formula the
generator uses (scale * X + offset).

  stacks   600k points: 8 stacks of 501 points at one position each and 4
           clusters of 4000 points jittered by -1..+1 unit per axis.  Ties of
           equal distance far longer than a wavefront (the libstdc++ order of
           sampler_poisson.h:182-203 over long runs) and long runs of equal
           float keys with distinct doubles.
  dups     400k points: a stack of 15 000 points at one position and one of
           10 001 at another: buildHierarchy's duplicate path
           (indexer.cpp:1002-1098), which keeps the first #distinct records.
  spacing  300k points in a 512 m cube (two anchor points), plus flat square
           lattices of pitch 1 m and 0.5 m in the air: neighbours exactly
           one spacing apart at levels 2 and 3 (4 m / 2^level), where the
           reference's `dd < squaredSpacing` accepts both.
  sphere   250k points on a sphere of radius 200 m (50 root spacings) around
           the centre of a 512 m cube: every accepted point of the root lies
           in the shell the Poisson scan walks, so its 10 001-check limit
           (sampler_poisson.h:172-175) binds for most root candidates.
"""
import struct

import numpy as np

from helpers import make_las

HDR_OFFSET_TO_POINTS = 96
HDR_RECORD_LENGTH = 105
HDR_NUM_POINTS = 247
HDR_SCALE = 131
HDR_OFFSET = 155
HDR_MAX_X = 179  # max x, min x, max y, min y, max z, min z (doubles)

SCENES = {
    # name: (preset, points, seed)
    "stacks": (0, 600_000, 21),
    "dups": (0, 400_000, 22),
    "spacing": (0, 300_000, 23),
    "sphere": (0, 250_000, 24),
}


def _load(path):
    data = bytearray(open(path, "rb").read())
    off = struct.unpack_from("<I", data, HDR_OFFSET_TO_POINTS)[0]
    rl = struct.unpack_from("<H", data, HDR_RECORD_LENGTH)[0]
    n = struct.unpack_from("<Q", data, HDR_NUM_POINTS)[0]
    recs = np.frombuffer(bytes(data[off:off + n * rl]), np.uint8).reshape(n, rl).copy()
    return data, off, recs


def xyz_of(recs):
    return recs[:, :12].copy().view(np.int32).reshape(-1, 3)


def _set_xyz(recs, rows, xyz):
    recs[rows, :12] = np.ascontiguousarray(np.asarray(xyz, np.int32).reshape(-1, 3)).view(np.uint8).reshape(-1, 12)


def _store(path, data, off, recs):
    data[off:off + recs.size] = recs.tobytes()
    q = xyz_of(recs)
    mn, mx = q.min(axis=0), q.max(axis=0)
    for a in range(3):
        s = struct.unpack_from("<d", data, HDR_SCALE + 8 * a)[0]
        o = struct.unpack_from("<d", data, HDR_OFFSET + 8 * a)[0]
        struct.pack_into("<d", data, HDR_MAX_X + 16 * a, s * int(mx[a]) + o)
        struct.pack_into("<d", data, HDR_MAX_X + 16 * a + 8, s * int(mn[a]) + o)
    with open(path, "wb") as f:
        f.write(bytes(data))


def _stacks(recs, rng):
    n = len(recs)
    q = xyz_of(recs)
    rows = rng.permutation(np.arange(2, n))
    k = 0
    for _ in range(8):  # exact stacks: 500 more points on an existing one
        anchor = rows[k]
        k += 1
        _set_xyz(recs, rows[k:k + 500], np.repeat(q[anchor][None, :], 500, axis=0))
        k += 500
    lo, hi = q.min(axis=0) + 10, q.max(axis=0) - 10
    for _ in range(4):  # near-duplicate clusters: -1..+1 unit per axis
        c = lo + (rng.random(3) * (hi - lo)).astype(np.int64)
        jit = rng.integers(-1, 2, size=(4000, 3))
        _set_xyz(recs, rows[k:k + 4000], c[None, :] + jit)
        k += 4000


def _dups(recs, rng):
    n = len(recs)
    q = xyz_of(recs)
    rows = rng.permutation(np.arange(2, n))
    anchors = rows[:2]
    _set_xyz(recs, rows[2:15001], np.repeat(q[anchors[0]][None, :], 14999, axis=0))
    _set_xyz(recs, rows[15001:25001], np.repeat(q[anchors[1]][None, :], 10000, axis=0))


def _spacing(recs, rng):
    q = xyz_of(recs)
    zmin = int(q[:, 2].min())
    # the cube: 512 m along x and y (anchors in records 0 and 1)
    _set_xyz(recs, [0, 1], [[0, 0, zmin], [512000, 512000, zmin]])
    rows = rng.permutation(np.arange(2, len(recs)))
    k = 0
    for pitch, z, x0 in ((1000, 300000, 100000), (500, 350000, 140000)):
        a, b = np.meshgrid(np.arange(30), np.arange(30), indexing="ij")
        pts = np.stack([x0 + pitch * a.ravel(), x0 + pitch * b.ravel(), np.full(900, z)], axis=1)
        _set_xyz(recs, rows[k:k + 900], pts)
        k += 900


def _sphere(recs, rng):
    n = len(recs)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pts = np.rint((256.0 + 200.0 * v) * 1000.0).astype(np.int64)
    _set_xyz(recs, np.arange(n), pts)
    _set_xyz(recs, [0, 1], [[0, 0, 0], [512000, 512000, 512000]])


EDITS = {"stacks": _stacks, "dups": _dups, "spacing": _spacing, "sphere": _sphere}


def make_scene(name, path):
    """Writes scene `name` to `path`; returns its point count."""
    preset, n, seed = SCENES[name]
    make_las(path, preset, n, seed)
    data, off, recs = _load(path)
    EDITS[name](recs, np.random.default_rng(seed))
    _store(path, data, off, recs)
    return n
