"""Structural checks of an oracle conversion, read the way the viewer reads it."""
import struct

from helpers import make_las, oracle_convert
from potree_format import decode_hierarchy, load_output


def test_hierarchy_and_octree_are_consistent(tmp_path):
    las = tmp_path / "a.las"
    make_las(las, 0, 400_000, 21)
    out = tmp_path / "o"
    oracle_convert(las, out, "poisson")
    meta_bytes, meta, hier, octree = load_output(out)
    nodes = decode_hierarchy(hier, meta["hierarchy"]["firstChunkSize"])
    assert "r" in nodes
    total = 0
    for name, n in nodes.items():
        assert n["type"] in (0, 1, 2)
        size = n.get("byteSize", 0)
        off = n.get("byteOffset", 0)
        assert off + size <= len(octree)
        assert size == 27 * n["numPoints"]
        total += n["numPoints"]
        for c in range(8):
            if n["childMask"] & (1 << c):
                assert name + str(c) in nodes
    # Poisson keeps every point exactly once
    assert total == 400_000 == meta["points"]
    assert len(octree) == 27 * 400_000
    assert meta["spacing"] > 0 and meta["hierarchy"]["stepSize"] == 4
