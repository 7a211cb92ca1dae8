"""Subtrees for the per-subtree sampling entry point (Sampler::sample).

Each case is a flattened subtree (node 0 = the node handed to sample(),
parents before children) built from generated pre-chunked records
(potreeconverter_amd.ChunkGen, 34-byte records, scale 0.001, offset 0, rgb
at byte 21): records are split into the eight children by the box centre
(x -> 4, y -> 2, z -> 1) until a node holds at most `cap` records, as
buildHierarchy would leave a chunk's subtree.

  chunk    one chunk's subtree below its level-4 root (phase 1 of the
           indexer, Converter/src/indexer.cpp:1638-1707)
  upper    a level-2 node above nine level-4 chunk roots that were sampled
           before (holders the traversal does not enter, phase 2)
  sparse   a level-5 node with leaves of one to three records next to a
           dense leaf: leaves emptied by their parent are discarded
"""
import struct

import numpy as np

import potreeconverter_amd as pc

BPP = 34
RGB = 21
SCALE = 0.001
CUBE = 2048.0
BASE_SPACING = CUBE / 128.0

CASES = {
    # name: (kind, ChunkGen arguments, cap)
    "chunk": ("chunk", dict(num_chunks=8, median=30_000, sigma=0.3, min_count=100, max_count=100_000, seed=5), 2_000),
    "upper": ("upper", dict(num_chunks=4096, median=3_000, sigma=0.3, min_count=100, max_count=100_000, seed=6), 0),
    "sparse": ("sparse", dict(num_chunks=8, median=5_000, sigma=0.1, min_count=100, max_count=100_000, seed=8), 0),
}


class Subtree:
    def __init__(self):
        self.children = []   # 8 per node
        self.bounds = []     # min xyz, max xyz per node
        self.levels = []
        self.sampled = []
        self.records = []    # bytes per node

    def add(self, mn, mx, level, parent=None, digit=None, recs=b"", sampled=0):
        i = len(self.levels)
        self.children.append([-1] * 8)
        self.bounds.append(list(mn) + list(mx))
        self.levels.append(level)
        self.sampled.append(sampled)
        self.records.append(recs)
        if parent is not None:
            self.children[parent][digit] = i
        return i

    @property
    def counts(self):
        return [len(r) // BPP for r in self.records]

    def input_file(self, path):
        """The input of oracle/_ref/ref_sampler (see oracle/ref_sampler.cpp)."""
        with open(path, "wb") as f:
            f.write(struct.pack("<iii", len(self.levels), BPP, RGB))
            f.write(struct.pack("<7d", SCALE, SCALE, SCALE, 0.0, 0.0, 0.0, BASE_SPACING))
            for i in range(len(self.levels)):
                f.write(struct.pack("<8i", *self.children[i]))
                f.write(struct.pack("<6d", *self.bounds[i]))
                f.write(struct.pack("<iiq", self.levels[i], self.sampled[i], self.counts[i]))
            for r in self.records:
                f.write(r)


def _xyz(recs):
    return np.frombuffer(recs, np.uint8).reshape(-1, BPP)[:, :12].copy().view(np.int32) * SCALE


def _split(t, node, recs, cap):
    """Distributes `recs` below `node` (a node without records yet)."""
    n = len(recs) // BPP
    if n <= cap or t.levels[node] >= 20:
        t.records[node] = recs
        return
    mn = np.array(t.bounds[node][:3])
    mx = np.array(t.bounds[node][3:])
    c = (mn + mx) / 2
    p = _xyz(recs)
    digit = ((p[:, 0] >= c[0]).astype(int) << 2) | ((p[:, 1] >= c[1]).astype(int) << 1) | (p[:, 2] >= c[2])
    arr = np.frombuffer(recs, np.uint8).reshape(-1, BPP)
    for d in range(8):
        sel = np.nonzero(digit == d)[0]
        if len(sel) == 0:
            continue
        bits = np.array([(d >> 2) & 1, (d >> 1) & 1, d & 1])
        half = (mx - mn) / 2
        cmn = mn + half * bits
        ch = t.add(cmn, cmn + half, t.levels[node] + 1, node, d)
        _split(t, ch, arr[sel].tobytes(), cap)


def _cell_box(name):
    """Box of a chunk "rABCD" in the cube."""
    mn, size = np.zeros(3), CUBE
    for d in name[1:]:
        d = int(d)
        size /= 2
        mn = mn + size * np.array([(d >> 2) & 1, (d >> 1) & 1, d & 1])
    return mn, mn + size


def make_subtree(name):
    kind, args, cap = CASES[name]
    g = pc.ChunkGen(**args)
    t = Subtree()
    if kind == "chunk":
        k = 3
        mn, mx = _cell_box(g.names[k])
        root = t.add(mn, mx, 4)
        _split(t, root, g.records(k), cap)
    elif kind == "upper":
        # level-2 node "r00": chunks r00AB; nine of its level-4 cells hold
        # every third record of their chunk (points a chunk root kept)
        mn, mx = _cell_box("r00")
        root = t.add(mn, mx, 2)
        picks = ["r0000", "r0001", "r0007", "r0012", "r0033", "r0050", "r0056", "r0077", "r0070"]
        mids = {}
        for nm in sorted(picks):
            k = g.names.index(nm)
            mid = nm[:4]
            if mid not in mids:
                a, b = _cell_box(mid)
                mids[mid] = t.add(a, b, 3, root, int(mid[3]))
            a, b = _cell_box(nm)
            recs = np.frombuffer(g.records(k), np.uint8).reshape(-1, BPP)[::3].tobytes()
            t.add(a, b, 4, mids[mid], int(nm[4]), recs, sampled=1)
    elif kind == "sparse":
        k = 5
        mn, mx = _cell_box(g.names[k])
        half = (mx - mn) / 2
        mn = mn + half * np.array([0, 0, 1])  # octant 1 of the chunk's cell: the terrain's upper half
        root = t.add(mn, mn + half, 5)
        arr = np.frombuffer(g.records(k), np.uint8).reshape(-1, BPP)
        p = _xyz(arr.tobytes())
        pts = arr[np.all((p >= mn) & (p < mn + half), axis=1)]
        q = half / 2
        for d in range(8):
            bits = np.array([(d >> 2) & 1, (d >> 1) & 1, d & 1])
            cmn = mn + q * bits
            p = _xyz(pts.tobytes())
            sel = np.nonzero(np.all((p >= cmn) & (p < cmn + q), axis=1))[0]
            # leaf 0 keeps its records, the others one to three of theirs
            recs = pts[sel].tobytes() if d == 0 else pts[sel[: 1 + d % 3]].tobytes()
            if len(recs):
                t.add(cmn, cmn + q, 6, root, d, recs)
    return t


def parse_reference_output(path, n):
    """oracle/_ref/ref_sampler's output: per node (count, children, sampled,
    records) and the callback events."""
    data = open(path, "rb").read()
    o = 0
    nodes = []
    for _ in range(n):
        cnt = struct.unpack_from("<q", data, o)[0]
        ch = list(struct.unpack_from("<8i", data, o + 8))
        sm = struct.unpack_from("<i", data, o + 40)[0]
        o += 44
        nodes.append((cnt, ch, sm, data[o:o + cnt * BPP]))
        o += cnt * BPP
    ne = struct.unpack_from("<i", data, o)[0]
    ev = [list(struct.unpack_from("<2i", data, o + 4 + 8 * e)) for e in range(ne)]
    return nodes, ev
