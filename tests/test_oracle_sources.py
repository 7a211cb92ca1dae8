"""doChunking over several sources with the --attributes filter, in the oracle.

The reference chunker cannot be built here, so the multi-source rules are
pinned by restating them independently in Python:
 - computeOutputAttributes(sources, requestedAttributes)
   (PotreeConverter.h:195-315): the union of the per-format attribute lists
   by name in source order, position first then the requested names that
   exist, the smallest LAS scale and the headers' bounds
   (computeScaleOffset, :29-76);
 - each chunk holds its points of the first source, then of the second, in
   input order (the canonical order of the reference's task pool), every
   point requantized against the common offset and scale;
 - a caller-given box (computeStats, main.cpp:215-284) is the cube of
   chunks/metadata.json; the output offset stays the headers' minimum;
 - extra-bytes attributes that would be written outside their output
   attribute (chunker_countsort_laszip.cpp:579-582) are refused.
"""
import json
import os
import struct

import numpy as np

from helpers import make_las, oracle, oracle_convert_sources, read

STD = {  # PotreeConverter.h:136-186: (name, size) per point format
    2: ["position:12", "intensity:2", "return number:1", "number of returns:1", "classification:1",
        "scan angle rank:1", "user data:1", "point source id:2", "rgb:6"],
    3: ["position:12", "intensity:2", "return number:1", "number of returns:1", "classification:1",
        "scan angle rank:1", "user data:1", "point source id:2", "gps-time:8", "rgb:6"],
}
EXTRA = ["gauss:4", "uniform16:2", "uniform8:1"]


def header(path):
    raw = read(path)
    scale = struct.unpack_from("<3d", raw, 131)
    mx = [struct.unpack_from("<d", raw, o)[0] for o in (179, 195, 211)]
    mn = [struct.unpack_from("<d", raw, o)[0] for o in (187, 203, 219)]
    return raw, np.array(scale), np.array(mn), np.array(mx)


def expected_attributes(formats_extra, requested):
    full = []
    for fmt, eb in formats_extra:
        for a in STD[fmt] + (EXTRA if eb else []):
            if a.split(":")[0] not in [b.split(":")[0] for b in full]:
                full.append(a)
    if requested:
        out = []
        for nm in ["position"] + requested:
            out += [a for a in full if a.split(":")[0] == nm][:1]
        return out
    return full


def meta_attributes(outdir):
    m = json.load(open(os.path.join(outdir, "chunks", "metadata.json")))
    return m, [f"{a['name']}:{a['size']}" for a in m["attributes"]]


def test_union_filter_scale_and_offset(tmp_path):
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 60_000, 1, extra_bytes=True, point_format=2)
    make_las(b, 2, 50_000, 2, extra_bytes=False, point_format=3)
    for requested in ([], ["rgb", "intensity", "nothing", "gps-time"]):
        out = tmp_path / f"o{len(requested)}"
        oracle_convert_sources([a, b], out, "poisson", attributes=requested, chunk_only=True)
        m, names = meta_attributes(out)
        assert names == expected_attributes([(2, True), (3, False)], requested)
        _, sa, mna, mxa = header(a)
        _, sb, mnb, mxb = header(b)
        mn, mx = np.minimum(mna, mnb), np.maximum(mxa, mxb)
        scale = np.maximum(np.minimum(sa, sb), (mx - mn) / 2.0 ** 30)
        assert m["scale"] == list(scale) and m["offset"] == list(mn)
        cube = (mx - mn).max()
        assert m["min"] == list(mn) and np.allclose(m["max"], mn + cube, rtol=0, atol=0)


def test_extra_bytes_that_would_move_are_refused(tmp_path):
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 30_000, 1, extra_bytes=False, point_format=3)
    make_las(b, 0, 30_000, 2, extra_bytes=True, point_format=2)
    # b's extra bytes: offset 27 in its own layout, 35 in the union (after gps-time)
    assert oracle_convert_sources([a, b], tmp_path / "o", "poisson", chunk_only=True, check=False) != 0
    assert "outside" in oracle().orc_last_error().decode()
    # a filter that drops a standard attribute before them moves them too
    assert oracle_convert_sources([b], tmp_path / "p", "poisson", attributes=["gauss"], chunk_only=True,
                                  check=False) != 0
    # keeping every standard attribute before them keeps them in place
    oracle_convert_sources([b, a], tmp_path / "q", "poisson", chunk_only=True)


def test_chunks_hold_sources_in_order(tmp_path):
    paths = [tmp_path / f"s{i}.las" for i in range(3)]
    for i, p in enumerate(paths):
        make_las(p, 0, 40_000 + 5_000 * i, 10 + i)
    out = tmp_path / "o"
    oracle_convert_sources(paths, out, "poisson", chunk_only=True)
    m, _ = meta_attributes(out)
    ps, po = np.array(m["scale"]), np.array(m["offset"])
    # every source's points requantized against the common offset and scale
    keys = {}
    total = 0
    for si, p in enumerate(paths):
        raw, scale, _, _ = header(p)
        off = struct.unpack_from("<I", raw, 96)[0]
        n = struct.unpack_from("<Q", raw, 247)[0]
        rl = struct.unpack_from("<H", raw, 105)[0]
        lo = np.array(struct.unpack_from("<3d", raw, 155))
        xyz = np.frombuffer(raw[off:off + n * rl], np.uint8).reshape(n, rl)[:, :12].copy().view(np.int32)
        q = np.trunc((xyz * scale + lo - po) / ps).astype(np.int64)
        for i in range(n):
            keys.setdefault(tuple(q[i]), []).append((si, i))
        total += n
    seen = 0
    for f in sorted(os.listdir(out / "chunks")):
        if not f.endswith(".bin"):
            continue
        recs = np.frombuffer(read(out / "chunks" / f), np.uint8).reshape(-1, 27)
        pos = recs[:, :12].copy().view(np.int32).astype(np.int64)
        seen += len(recs)
        order = [keys[tuple(p)][0] for p in pos]
        for k in range(1, len(order)):  # (source, index) ascending inside a chunk
            if order[k] < order[k - 1]:
                # a position shared by several points: any of them may match
                assert len(keys[tuple(pos[k])]) > 1 or len(keys[tuple(pos[k - 1])]) > 1
    assert seen == total


def test_caller_box(tmp_path):
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 30_000, 3)
    make_las(b, 0, 30_000, 4)
    _, _, mna, mxa = header(a)
    _, _, mnb, mxb = header(b)
    mn = np.minimum(mna, mnb) - 7.0
    cube = (np.maximum(mxa, mxb) - mn).max() + 11.0
    out = tmp_path / "o"
    oracle_convert_sources([a, b], out, "poisson", cube_min=list(mn), cube_max=list(mn + cube), chunk_only=True)
    m, _ = meta_attributes(out)
    assert m["min"] == list(mn) and m["max"] == list(mn + cube)
    assert m["offset"] == list(np.minimum(mna, mnb))


def test_orphaned_hierarchy_chunk_is_an_error(tmp_path):
    """A written node whose hierarchy chunk root (r0000) was unlinked unwritten
    by Poisson-average: the reference indexer crashes on it
    (HierarchyBuilder.h:317-326); the oracle stops with 123."""
    a, b = tmp_path / "a.las", tmp_path / "b.las"
    make_las(a, 0, 300_000, 4, True, 2)
    make_las(b, 2, 200_000, 5, False, 3)
    assert oracle_convert_sources([a, b], tmp_path / "o", "poisson_average", check=False) == 123
    assert "didn't find proxy node r0000" in oracle().orc_last_error().decode()
